// Host side of libptg: context, device memory, the generation loop that
// replaces the per-rank collide() map (src/main.rs:1118-1135), and the C ABI
// declared in include/ptg.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dlfcn.h>
#include <functional>
#include <map>
#include <string>
#include <vector>

#include "kern_lma.h"
#include "dev_bw.h"
#include "kern_misc.h"
#include "mlp_host.h"

using namespace ptg;

namespace {

constexpr double PI_H = 3.14159265358979323846;
constexpr double EPS0_SI = 8.854188e-12;  // vacuum permittivity, 7 digits as src/constants.rs:8

// Device particle buffer with capacity management.
struct DevBuf {
    void* base = nullptr;
    size_t cap = 0;
    PBuf v{};
};

hipError_t alloc_buf(DevBuf& b, size_t cap) {
    if (cap < 64) cap = 64;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, cap * PBUF_BYTES_PER);
    if (e != hipSuccess) return e;
    char* c = static_cast<char*>(p);
    PBuf v;
    for (int k = 0; k < P_NDOUBLE; ++k) { v.d[k] = reinterpret_cast<double*>(c); c += cap * 8; }
    v.id = reinterpret_cast<uint64_t*>(c); c += cap * 8;
    v.parent_id = reinterpret_cast<uint64_t*>(c); c += cap * 8;
    v.stream = reinterpret_cast<uint64_t*>(c); c += cap * 8;
    v.shower = reinterpret_cast<uint32_t*>(c); c += cap * 4;
    v.meta = reinterpret_cast<uint32_t*>(c); c += cap * 4;
    v.nev = reinterpret_cast<uint32_t*>(c);
    if (b.base) (void)hipFree(b.base);
    b.base = p;
    b.cap = cap;
    b.v = v;
    return hipSuccess;
}

// Grow keeping the first `keep` records.
hipError_t grow_buf(DevBuf& b, size_t cap, size_t keep, hipStream_t s) {
    if (cap <= b.cap) return hipSuccess;
    DevBuf nb;
    hipError_t e = alloc_buf(nb, cap);
    if (e != hipSuccess) return e;
    if (b.base && keep) {
        for (int k = 0; k < P_NDOUBLE; ++k) (void)hipMemcpyAsync(nb.v.d[k], b.v.d[k], keep * 8, hipMemcpyDeviceToDevice, s);
        (void)hipMemcpyAsync(nb.v.id, b.v.id, keep * 8, hipMemcpyDeviceToDevice, s);
        (void)hipMemcpyAsync(nb.v.parent_id, b.v.parent_id, keep * 8, hipMemcpyDeviceToDevice, s);
        (void)hipMemcpyAsync(nb.v.stream, b.v.stream, keep * 8, hipMemcpyDeviceToDevice, s);
        (void)hipMemcpyAsync(nb.v.shower, b.v.shower, keep * 4, hipMemcpyDeviceToDevice, s);
        (void)hipMemcpyAsync(nb.v.meta, b.v.meta, keep * 4, hipMemcpyDeviceToDevice, s);
        (void)hipMemcpyAsync(nb.v.nev, b.v.nev, keep * 4, hipMemcpyDeviceToDevice, s);
        (void)hipStreamSynchronize(s);
    }
    if (b.base) (void)hipFree(b.base);
    b = nb;
    return hipSuccess;
}

void free_buf(DevBuf& b) {
    if (b.base) (void)hipFree(b.base);
    b = DevBuf{};
}

// Device scratch of ptg_histogram_noalloc: words [0, nbins) are the bins and
// [nbins] the total weight (doubles), [PTG_NOALLOC_MAX_BINS + 1] and [+ 2] the
// order keys of the range search. Each GPU holds its own copy until the process
// ends; host code reaches it only through the address hipGetSymbolAddress gives
// on that device (ptg_ctx::d_noalloc), never through this name.
static __device__ __align__(256) unsigned long long scratch[PTG_NOALLOC_MAX_BINS + 3];

std::string default_table_path() {
    Dl_info info;
    if (dladdr(reinterpret_cast<void*>(&default_table_path), &info) && info.dli_fname) {
        std::string p(info.dli_fname);
        size_t k = p.rfind('/');
        std::string dir = k == std::string::npos ? "." : p.substr(0, k);
        return dir + "/tables/ptg_tables.bin";
    }
    return "tables/ptg_tables.bin";
}

}  // namespace

struct ptg_ctx {
    int device = 0;
    int n_cu = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::map<std::string, std::vector<double>> tables;
    double* d_lcfa = nullptr;
    LcfaScalars sc{};
    double* d_lma = nullptr;
    float* d_lma_f = nullptr;  // FP32 copy of the LMA pack (mixed precision)
    LmaScalars lsc{};
    double* d_lp_ceil = nullptr;    // tabulated LP emission ceiling (lp_ceil::); null if absent
    double* d_bw = nullptr;         // LP Breit-Wheeler pack (dev_bw.h); null if its tables are absent
    BwScalars bsc{};
    double* d_cp = nullptr;         // CP nonlinear Compton pack (dev_cp.h); null if its tables are absent
    CpScalars csc{};
    double lma_bandwidth = 0.0, pol_angle = 0.0;
    DevBuf evq, lq[2], phq, phq2, pairq;
    double* d_ev_extra = nullptr;  // ev_k0, ev_a
    size_t ev_extra_cap = 0;
    LaserDev laser{};
    bool have_laser = false;
    DevBuf prim, q[2], out, inter, tmp;
    double* d_n0 = nullptr;
    size_t n0_cap = 0;
    size_t n_prim = 0;
    uint32_t* d_shower_count = nullptr;  // final-state particles per shower (multiplicity selection)
    size_t shower_cap = 0;
    // secondary-queue records per input particle seen so far (lepton, photon stage):
    // sizes the queue of the next transport launch so that one launch suffices
    double next_ratio[2] = {2.0, 2.0};
    unsigned long long* d_ctr = nullptr;
    TransportArgs* d_args = nullptr;
    uint64_t n_out = 0, n_inter = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double* d_hist = nullptr;               // ptg_histogram's bins and total, grown as needed
    size_t hist_cap = 0;
    unsigned long long* d_noalloc = nullptr;  // this device's copy of `scratch` (ptg_histogram_noalloc)
    // learned sampler, one per polarization (0 linear, 1 circular)
    MlpNet surr[2];
    bool surr_loaded[2] = {false, false};
    bool surr_valid[2] = {false, false};
    double surr_box[2][6] = {};    // a_min, a_max, eta_min, eta_max, a_use_min, eta_use_min
    uint32_t* d_surr_list = nullptr;  // row -> event, and the exact sampler's list
    // a batch of collisions (ptg_set_batch): lasers, and each primary's collision
    LaserDev* d_lasers = nullptr;
    uint8_t* d_coll = nullptr;
    int spectrum = 0, n_cdf = 0;        // the next beams' energy spectrum (ptg_set_beam_spectrum)
    double* d_cdf = nullptr;
    double brem_min = 0.0, brem_max = 0.0;
    size_t coll_cap = 0;
    int n_coll = 0;
    double batch_a0_max = 0.0;
    // launch shape of the lepton step kernel (ptg_autotune)
    int step_block = 256, step_grid_per_cu = 4;
    uint32_t* d_surr_fb = nullptr;
    uint32_t* d_surr_now = nullptr;     // events below the surrogate's amplitude range
    uint32_t* d_surr_perm = nullptr;
    unsigned int* d_surr_bucket = nullptr;  // [2][32]
    size_t surr_list_cap = 0;
    // emission events the learned sampler leaves to the exact kernels, held until
    // the lepton phase drains (one exact pass per phase instead of one per generation)
    DevBuf defq;
    double* d_def_extra = nullptr;  // kappa^0, sqrt(<a^2>)
    size_t def_extra_cap = 0, n_def = 0;
    // LMA generations: PTG_LAUNCH_GRAPH runs them in captured sequences (launch
    // arguments of the two queue parities in d_gargs[2]); PTG_LAUNCH_EAGER reads
    // the counts back after every launch
    int launch_mode = PTG_LAUNCH_GRAPH;
    TransportArgs* d_gargs = nullptr;
    hipGraphExec_t gexec = nullptr;
    uint64_t graph_launches = 0;  // of the current ptg_run
    uint32_t graph_stalls = 0;
};

namespace {

int fail(ptg_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}
#define HIPCHK(ctx, expr)                                                                         \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail(ctx, e_ == hipErrorOutOfMemory ? PTG_E_OOM : PTG_E_HIP,                   \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                       \
    } while (0)

// Every copy and fill goes through the context's stream: the stream is created
// non-blocking, so the legacy-stream hipMemcpy/hipMemset would not be ordered
// after kernels already queued on it. Copies are synchronous for the caller.
static hipError_t scopy(ptg_ctx* c, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, c->stream);
    return e != hipSuccess ? e : hipStreamSynchronize(c->stream);
}
static hipError_t sfill(ptg_ctx* c, void* dst, int value, size_t bytes) {
    return hipMemsetAsync(dst, value, bytes, c->stream);
}

bool load_blob(const std::string& path, std::map<std::string, std::vector<double>>& out, std::string& err) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) { err = "cannot open table blob " + path; return false; }
    std::fseek(f, 0, SEEK_END);
    long len = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    std::vector<char> raw(len > 0 ? len : 0);
    bool ok = len > 16 && std::fread(raw.data(), 1, len, f) == size_t(len);
    std::fclose(f);
    if (!ok || std::memcmp(raw.data(), "PTGTBL01", 8) != 0) { err = "malformed table blob " + path; return false; }
    uint32_t n;
    std::memcpy(&n, raw.data() + 8, 4);
    const size_t rec = 96 + 8 + 8 + 4 + 16 + 4;
    size_t base = 16 + size_t(n) * rec;
    for (uint32_t i = 0; i < n; ++i) {
        const char* r = raw.data() + 16 + i * rec;
        char name[97] = {0};
        std::memcpy(name, r, 96);
        uint64_t off, cnt;
        std::memcpy(&off, r + 96, 8);
        std::memcpy(&cnt, r + 104, 8);
        if (base + (off + cnt) * 8 > size_t(len)) { err = "table blob truncated"; return false; }
        std::vector<double> v(cnt);
        std::memcpy(v.data(), raw.data() + base + off * 8, cnt * 8);
        out[name] = std::move(v);
    }
    return true;
}

bool build_lcfa_pack(ptg_ctx* c, std::vector<double>& pack) {
    using namespace lcfa_pack;
    pack.assign(SIZE, 0.0);
    auto put = [&](int off, const char* name, size_t expect) -> bool {
        auto it = c->tables.find(name);
        if (it == c->tables.end() || it->second.size() != expect) {
            c->err = std::string("table ") + name + " missing or wrong size";
            return false;
        }
        std::copy(it->second.begin(), it->second.end(), pack.begin() + off);
        return true;
    };
    const char* PE = "lcfa/photon_emission/tables:";
    std::string pe(PE);
    bool ok = put(LN_H_CHI, (pe + "LN_H_CHI_TABLE").c_str(), 82) && put(LN_T_CHI, "lcfa/pair_creation/tables:LN_T_CHI_TABLE", 243) &&
              put(QUANTUM_CDF, (pe + "QUANTUM_CDF").c_str(), 41 * 64) && put(Y_CDF, (pe + "Y_CDF").c_str(), 31 * 64) &&
              put(Y_INF, (pe + "Y_INFINITE_DELTA_CDF").c_str(), 64) && put(HIGH_CHI, (pe + "HIGH_CHI_CDF").c_str(), 64) &&
              put(HIGH_CHI_PEAK, (pe + "HIGH_CHI_PEAK_CDF").c_str(), 64) &&
              put(CLASSICAL, (pe + "CLASSICAL_SPECTRUM_TABLE").c_str(), 82) &&
              put(GAUNT, "lcfa/photon_emission/mod:GAUNT_FACTOR_TABLE", 50) && put(GL_NODES, "quadrature/tables:GL_NODES", 32) &&
              put(GL_WEIGHTS, "quadrature/tables:GL_WEIGHTS", 32) &&
              put(AIRY_SERIES, "special_functions/airy:AIRY_SMALL_X", 40) &&
              put(AIRYP_SERIES, "special_functions/airy:AIRY_PRIME_SMALL_X", 40) &&
              put(AIRY_N40, "special_functions/airy:GL_m1_6_NODES_40", 40) && put(AIRY_W40, "special_functions/airy:GL_m1_6_WEIGHTS_40", 40) &&
              put(AIRY_N16, "special_functions/airy:GL_m1_6_NODES_16", 16) && put(AIRY_W16, "special_functions/airy:GL_m1_6_WEIGHTS_16", 16) &&
              put(AIRY_N4, "special_functions/airy:GL_m1_6_NODES_4", 4) && put(AIRY_W4, "special_functions/airy:GL_m1_6_WEIGHTS_4", 4) &&
              put(AIRYP_N32, "special_functions/airy:GL_p1_6_NODES_32", 32) && put(AIRYP_W32, "special_functions/airy:GL_p1_6_WEIGHTS_32", 32) &&
              put(AIRYP_N16, "special_functions/airy:GL_p1_6_NODES_16", 16) && put(AIRYP_W16, "special_functions/airy:GL_p1_6_WEIGHTS_16", 16) &&
              put(AIRYP_N8, "special_functions/airy:GL_p1_6_NODES_8", 8) && put(AIRYP_W8, "special_functions/airy:GL_p1_6_WEIGHTS_8", 8) &&
              put(AIRYP_N4, "special_functions/airy:GL_p1_6_NODES_4", 4) && put(AIRYP_W4, "special_functions/airy:GL_p1_6_WEIGHTS_4", 4) &&
              put(AIRY_CHEB, "special_functions/airy:CHEB_GL_SUMS", 190);
    if (!ok) return false;
    // node-only factors of the pair spectrum quadrature (lcfa/pair_creation/mod.rs:80-89)
    for (int k = 0; k < 32; ++k) {
        double t = pack[GL_NODES + k], w = pack[GL_WEIGHTS + k];
        double* q = pack.data() + PAIR_QN + 5 * k;
        q[0] = t;
        q[1] = w;
        q[2] = std::cosh(t);
        q[3] = std::cosh(2.0 * t / 3.0);
        q[4] = std::cosh(t / 3.0) / std::cosh(t);
    }
    auto sc1 = [&](const char* name) { return c->tables[name][0]; };
    c->sc.delta_ln_chi_h = sc1("lcfa/photon_emission/tables:DELTA_LN_CHI");
    c->sc.ln_chi_min = sc1("lcfa/photon_emission/tables:LN_CHI_MIN");
    c->sc.ln_chi_step = sc1("lcfa/photon_emission/tables:LN_CHI_STEP");
    c->sc.ln_delta_min = sc1("lcfa/photon_emission/tables:LN_DELTA_MIN");
    c->sc.ln_delta_step = sc1("lcfa/photon_emission/tables:LN_DELTA_STEP");
    c->sc.delta_ln_chi_t = sc1("lcfa/pair_creation/tables:DELTA_LN_CHI");
    return true;
}

bool build_lma_pack(ptg_ctx* c, std::vector<double>& pack) {
    pack.assign(lma_pack::SIZE, 0.0);
    auto need = [&](const char* name, size_t n) -> const std::vector<double>* {
        auto it = c->tables.find(name);
        if (it == c->tables.end() || it->second.size() != n) {
            c->err = std::string("table ") + name + " missing or wrong size";
            return nullptr;
        }
        return &it->second;
    };
    const auto* rate = need("nonlinear_compton/lp/rate_table:TABLE", lma_pack::RATE_COLS * lma_pack::RATE_ROWS);
    const auto* cdf = need("nonlinear_compton/lp/cdf_table:TABLE", size_t(lma_pack::CDF_N) * 32);
    const auto* rmin = need("nonlinear_compton/lp/rate_table:MIN", 2);
    const auto* rstep = need("nonlinear_compton/lp/rate_table:STEP", 2);
    const auto* cmin = need("nonlinear_compton/lp/cdf_table:MIN", 2);
    const auto* cstep = need("nonlinear_compton/lp/cdf_table:STEP", 2);
    const auto* cl_rate = need("nonlinear_compton/lp/classical/rate_table:TABLE", lma_pack::CL_N);
    const auto* cl_cdf = need("nonlinear_compton/lp/classical/cdf_table:TABLE", size_t(lma_pack::CL_N) * 32);
    const auto* cl_rmin = need("nonlinear_compton/lp/classical/rate_table:MIN", 1);
    const auto* cl_rstep = need("nonlinear_compton/lp/classical/rate_table:STEP", 1);
    const auto* cl_cmin = need("nonlinear_compton/lp/classical/cdf_table:MIN", 1);
    const auto* cl_cstep = need("nonlinear_compton/lp/classical/cdf_table:STEP", 1);
    if (!rate || !cdf || !rmin || !rstep || !cmin || !cstep || !cl_rate || !cl_cdf || !cl_rmin || !cl_rstep || !cl_cmin || !cl_cstep)
        return false;
    std::copy(rate->begin(), rate->end(), pack.begin() + lma_pack::RATE);
    std::copy(cdf->begin(), cdf->end(), pack.begin() + lma_pack::CDF);
    std::copy(cl_rate->begin(), cl_rate->end(), pack.begin() + lma_pack::CL_RATE);
    std::copy(cl_cdf->begin(), cl_cdf->end(), pack.begin() + lma_pack::CL_CDF);
    c->lsc.cl_rate_min = (*cl_rmin)[0];
    c->lsc.cl_rate_step = (*cl_rstep)[0];
    c->lsc.cl_cdf_min = (*cl_cmin)[0];
    c->lsc.cl_cdf_step = (*cl_cstep)[0];
    for (int k = 0; k < 2; ++k) {
        c->lsc.rate_min[k] = (*rmin)[k];
        c->lsc.rate_step[k] = (*rstep)[k];
        c->lsc.cdf_min[k] = (*cmin)[k];
        c->lsc.cdf_step[k] = (*cstep)[k];
    }
    return true;
}

// LP nonlinear Breit-Wheeler tables (dev_bw.h). The mid-range CDF is regenerated
// (tables/gen_bw_tables.py); without it, pair creation in a PlaneWave is unavailable.
bool build_bw_pack(ptg_ctx* c, std::vector<double>& pack) {
    using namespace bw_pack;
    pack.assign(SIZE, 0.0);
    auto get = [&](const char* name, size_t n) -> const std::vector<double>* {
        auto it = c->tables.find(name);
        return (it == c->tables.end() || it->second.size() != n) ? nullptr : &it->second;
    };
    const auto* hi = get("pair_creation/lp/tables/total:TABLE", size_t(HI_ROWS) * HI_COLS * 2);
    const auto* hic = get("pair_creation/lp/tables/cdf:TABLE", size_t(HI_CDF_N) * 48);
    const auto* mid = get("pair_creation/lp/tables/mid_range/total:TABLE", size_t(MID_ROWS) * MID_COLS * 2);
    const auto* midc = get("pair_creation/lp/tables/mid_range/cdf:TABLE", size_t(MID_CDF_N) * 48);
    auto one = [&](const char* name) { const auto* v = get(name, 1); return v ? (*v)[0] : NAN; };
    const auto* cmin = get("pair_creation/lp/tables/cdf:MIN", 2);
    const auto* cstep = get("pair_creation/lp/tables/cdf:STEP", 2);
    if (!hi || !hic || !mid || !midc || !cmin || !cstep) return false;
    std::copy(hi->begin(), hi->end(), pack.begin() + HI_TOTAL);
    std::copy(hic->begin(), hic->end(), pack.begin() + HI_CDF);
    std::copy(mid->begin(), mid->end(), pack.begin() + MID_TOTAL);
    std::copy(midc->begin(), midc->end(), pack.begin() + MID_CDF);
    BwScalars& b = c->bsc;
    b.hi_ln_min_a = one("pair_creation/lp/tables/total:LN_MIN_A");
    b.hi_ln_min_eta = one("pair_creation/lp/tables/total:LN_MIN_ETA");
    b.hi_ln_a_step = one("pair_creation/lp/tables/total:LN_A_STEP");
    b.hi_ln_eta_step = one("pair_creation/lp/tables/total:LN_ETA_STEP");
    for (int k = 0; k < 2; ++k) { b.hi_cdf_min[k] = (*cmin)[k]; b.hi_cdf_step[k] = (*cstep)[k]; }
    b.mid_ln_min_a = one("pair_creation/lp/tables/mid_range/total:LN_MIN_A");
    b.mid_ln_max_eta_prime = one("pair_creation/lp/tables/mid_range/total:LN_MAX_ETA_PRIME");
    b.mid_ln_min_eta_prime = one("pair_creation/lp/tables/mid_range/total:LN_MIN_ETA_PRIME");
    b.mid_ln_a_step = one("pair_creation/lp/tables/mid_range/total:LN_A_STEP");
    b.mid_eta_prime_density = one("pair_creation/lp/tables/mid_range/total:ETA_PRIME_DENSITY");
    return std::isfinite(b.hi_ln_min_a) && std::isfinite(b.mid_eta_prime_density);
}

bool build_cp_pack(ptg_ctx* c, std::vector<double>& pack) {
    using namespace cp_pack;
    pack.assign(SIZE, 0.0);
    auto get = [&](const char* name, size_t n) -> const std::vector<double>* {
        auto it = c->tables.find(name);
        return (it == c->tables.end() || it->second.size() != n) ? nullptr : &it->second;
    };
    const auto* rate = get("nonlinear_compton/cp/tables/total:TABLE", size_t(CDF_N));
    const auto* cdf = get("nonlinear_compton/cp/tables/cdf:TABLE", size_t(CDF_N) * 64);
    const auto* cl_rate = get("nonlinear_compton/cp/classical/rate_table:TABLE", size_t(CL_N));
    const auto* cl_cdf = get("nonlinear_compton/cp/classical/cdf_table:TABLE", size_t(CL_N) * 64);
    const auto* g16 = get("quadrature/tables:GAUSS_16_NODES", 16);
    const auto* g32n = get("quadrature/tables:GAUSS_32_NODES", 32);
    const auto* g32w = get("quadrature/tables:GAUSS_32_WEIGHTS", 32);
    const auto* bw = get("pair_creation/cp/total:TABLE", size_t(BW_COLS) * BW_ROWS * 2);
    const auto* mn = get("nonlinear_compton/cp/tables/total:MIN", 2);
    const auto* st = get("nonlinear_compton/cp/tables/total:STEP", 2);
    auto one = [&](const char* name) { const auto* v = get(name, 1); return v ? (*v)[0] : NAN; };
    if (!rate || !cdf || !cl_rate || !cl_cdf || !g16 || !g32n || !g32w || !mn || !st || !bw) return false;
    std::copy(bw->begin(), bw->end(), pack.begin() + BW_TOTAL);
    std::copy(rate->begin(), rate->end(), pack.begin() + RATE);
    std::copy(cdf->begin(), cdf->end(), pack.begin() + CDF);
    std::copy(cl_rate->begin(), cl_rate->end(), pack.begin() + CL_RATE);
    std::copy(cl_cdf->begin(), cl_cdf->end(), pack.begin() + CL_CDF);
    std::copy(g16->begin(), g16->end(), pack.begin() + G16N);
    std::copy(g32n->begin(), g32n->end(), pack.begin() + G32N);
    std::copy(g32w->begin(), g32w->end(), pack.begin() + G32W);
    CpScalars& q = c->csc;
    for (int k = 0; k < 2; ++k) { q.min[k] = (*mn)[k]; q.step[k] = (*st)[k]; }
    q.cl_rate_min = one("nonlinear_compton/cp/classical/rate_table:MIN");
    q.cl_rate_step = one("nonlinear_compton/cp/classical/rate_table:STEP");
    q.cl_cdf_min = one("nonlinear_compton/cp/classical/cdf_table:MIN");
    q.cl_cdf_step = one("nonlinear_compton/cp/classical/cdf_table:STEP");
    q.bw_ln_min_a = one("pair_creation/cp/total:LN_MIN_A");
    q.bw_ln_min_eta = one("pair_creation/cp/total:LN_MIN_ETA");
    q.bw_ln_a_step = one("pair_creation/cp/total:LN_A_STEP");
    q.bw_ln_eta_step = one("pair_creation/cp/total:LN_ETA_STEP");
    return std::isfinite(q.cl_rate_min) && std::isfinite(q.cl_cdf_step) && std::isfinite(q.bw_ln_min_a) && std::isfinite(q.bw_ln_eta_step);
}

// Laser constants (fast_focused_laser.rs:24-38,47-53,129-169; fast_plane_wave.rs:23-108;
// plane_wave.rs:27-63,176-192).
int make_laser(ptg_ctx* c, const ptg_laser_desc* d, LaserDev& L) {
    if (d->kind < 0 || d->kind > 3) return fail(c, PTG_E_ARG, "laser kind out of range");
    if (d->envelope < 0 || d->envelope > 2) return fail(c, PTG_E_ARG, "envelope out of range");
    if (!(d->a0 > 0.0) || !(d->wavelength > 0.0) || !(d->n_cycles > 0.0)) return fail(c, PTG_E_ARG, "a0, wavelength and n_cycles must be positive");
    if ((d->kind == PTG_FOCUSED_LASER || d->kind == PTG_FAST_FOCUSED_LASER) && !(d->waist > 0.0))
        return fail(c, PTG_E_ARG, "focused lasers need a positive waist");
    L = LaserDev{};
    L.kind = d->kind;
    L.envelope = d->envelope;
    L.circular = d->polarization == PTG_CIRCULAR;
    L.finite_bandwidth = d->finite_bandwidth;
    L.a0 = d->a0;
    L.k0 = 2.0 * PI_H / d->wavelength;
    L.omega = C_LIGHT * L.k0;
    double pol_angle = L.circular ? 0.0 : d->pol_angle;
    L.sin_pa = std::sin(pol_angle);
    L.cos_pa = std::cos(pol_angle);
    L.chirp_b = d->chirp_b;
    L.a_norm = E_CHARGE / (M_ELECTRON * C_LIGHT * L.omega);
    double multiplier = std::fmin(std::sqrt(3.0) / (5.0 * ALPHA * d->a0), 0.1);
    if (d->kind == PTG_FAST_FOCUSED_LASER) {
        double duration = d->n_cycles * d->wavelength / C_LIGHT;
        L.n_cycles = C_LIGHT * duration * L.k0 / (2.0 * PI_H);
        double zr = 0.5 * L.k0 * d->waist * d->waist;
        L.inv_waist = 1.0 / d->waist;
        L.inv_zr = 1.0 / zr;
        L.eps = d->waist / zr;
        L.E0 = M_ELECTRON * C_LIGHT * L.omega * d->a0 / E_CHARGE;
        L.B0 = L.E0 / C_LIGHT;
        double tau = L.omega * duration;
        L.gauss_c = 2.0 * LN2_D / (tau * tau);
        L.dt = multiplier / L.omega;
    } else if (d->kind == PTG_FAST_PLANE_WAVE) {
        L.n_cycles = d->n_cycles;
        L.amplitude = (M_ELECTRON * C_LIGHT_SQD * L.k0 * d->a0) / E_CHARGE;
        double chirp = 1.0 + 2.0 * d->chirp_b * PI_H * d->n_cycles;
        L.dt = multiplier / (C_LIGHT * L.k0 * chirp);
    } else if (d->kind == PTG_FOCUSED_LASER) {  // focused_laser.rs:25-70, 165-173: no chirp; n_cycles from the duration
        double duration = d->n_cycles * d->wavelength / C_LIGHT;
        L.n_cycles = C_LIGHT * duration * L.k0 / (2.0 * PI_H);
        L.chirp_b = 0.0;
        L.waist_sqd = d->waist * d->waist;
        L.zr = 0.5 * L.k0 * L.waist_sqd;
        double tau = L.omega * duration;
        L.tau_sqd = tau * tau;
        L.dt = (d->envelope == PTG_FLATTOP ? 0.2 : 1.0) / L.omega;
    } else {  // PlaneWave (LMA)
        L.n_cycles = d->n_cycles;
        L.dt = (d->envelope == PTG_FLATTOP ? 0.2 : 1.0) / (C_LIGHT * L.k0);
    }
    double nc = L.n_cycles;
    L.max_phase = d->envelope == PTG_COS_SQUARED ? PI_H * nc : (d->envelope == PTG_FLATTOP ? PI_H * (nc + 1.0) : 6.0 * PI_H * nc);
    return PTG_OK;
}

double ideal_initial_z(const ptg_laser_desc* d, const LaserDev& L) {
    double wl = 2.0 * PI_H / L.k0;
    double nc = L.n_cycles;
    return d->envelope == PTG_COS_SQUARED ? 0.5 * wl * nc : (d->envelope == PTG_FLATTOP ? 0.5 * wl * (nc + 1.0) : 2.0 * wl * nc);
}

int ensure_n0(ptg_ctx* c, size_t n) {
    if (n <= c->n0_cap) return PTG_OK;
    if (c->d_n0) (void)hipFree(c->d_n0);
    c->d_n0 = nullptr;
    HIPCHK(c, hipMalloc(&c->d_n0, 3 * n * sizeof(double)));
    c->n0_cap = n;
    return PTG_OK;
}

uint64_t key_from_seed(uint64_t seed) { return mix64(0x8658b90036b165ebULL + seed * 0x32f55cddaebae910ULL); }

}  // namespace

extern "C" {

int ptg_create(ptg_ctx** out, const char* table_path, int device) {
    if (!out) return PTG_E_ARG;
    *out = nullptr;
    ptg_ctx* c = new ptg_ctx;
    c->device = device;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) {
        c->err = "no HIP device " + std::to_string(device);
        *out = c;
        return PTG_E_HIP;
    }
    HIPCHK(c, hipSetDevice(device));
    {
        // this device's address of the no-allocation scratch array (once, at set-up)
        void* p = nullptr;
        HIPCHK(c, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        c->d_noalloc = static_cast<unsigned long long*>(p);
    }
    hipDeviceProp_t prop;
    HIPCHK(c, hipGetDeviceProperties(&prop, device));
    c->n_cu = prop.multiProcessorCount;
    std::string path = table_path ? table_path : default_table_path();
    if (!load_blob(path, c->tables, c->err)) { *out = c; return PTG_E_TABLES; }
    std::vector<double> pack;
    if (!build_lcfa_pack(c, pack)) { *out = c; return PTG_E_TABLES; }
    std::vector<double> lpack;
    if (!build_lma_pack(c, lpack)) { *out = c; return PTG_E_TABLES; }
    HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    std::vector<double> bpack;
    if (build_bw_pack(c, bpack)) {
        HIPCHK(c, hipMalloc(&c->d_bw, bpack.size() * sizeof(double)));
        HIPCHK(c, scopy(c, c->d_bw, bpack.data(), bpack.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    std::vector<double> cpack;
    if (build_cp_pack(c, cpack)) {
        HIPCHK(c, hipMalloc(&c->d_cp, cpack.size() * sizeof(double)));
        HIPCHK(c, scopy(c, c->d_cp, cpack.data(), cpack.size() * sizeof(double), hipMemcpyHostToDevice));
        launch_cp_cdf_logs(c->d_cp, c->stream);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    HIPCHK(c, hipMalloc(&c->d_lma, lpack.size() * sizeof(double)));
    HIPCHK(c, scopy(c, c->d_lma, lpack.data(), lpack.size() * sizeof(double), hipMemcpyHostToDevice));
    launch_lp_cdf_logs(c->d_lma, c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    {
        std::vector<float> lf(lpack.begin(), lpack.end());
        HIPCHK(c, hipMalloc(&c->d_lma_f, lf.size() * sizeof(float)));
        HIPCHK(c, scopy(c, c->d_lma_f, lf.data(), lf.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    {
        // the tabulated LP ceiling, when the table blob holds it (tables/gen_lp_ceiling.py)
        auto it = c->tables.find("nonlinear_compton/lp:CEILING_MAX");
        if (it != c->tables.end() && it->second.size() == size_t(lp_ceil::SIZE)) {
            HIPCHK(c, hipMalloc(&c->d_lp_ceil, it->second.size() * sizeof(double)));
            HIPCHK(c, scopy(c, c->d_lp_ceil, it->second.data(), it->second.size() * sizeof(double), hipMemcpyHostToDevice));
        }
    }
    HIPCHK(c, hipMalloc(&c->d_lcfa, pack.size() * sizeof(double)));
    HIPCHK(c, scopy(c, c->d_lcfa, pack.data(), pack.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, hipMalloc(&c->d_ctr, C_N * sizeof(unsigned long long)));
    HIPCHK(c, hipMalloc(&c->d_args, sizeof(TransportArgs)));
    HIPCHK(c, hipMalloc(&c->d_gargs, 2 * sizeof(TransportArgs)));
    HIPCHK(c, hipEventCreate(&c->ev0));
    HIPCHK(c, hipEventCreate(&c->ev1));
    *out = c;
    return PTG_OK;
}

void ptg_destroy(ptg_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_buf(c->prim); free_buf(c->q[0]); free_buf(c->q[1]); free_buf(c->out); free_buf(c->inter); free_buf(c->tmp);
    if (c->d_lcfa) (void)hipFree(c->d_lcfa);
    if (c->d_lma) (void)hipFree(c->d_lma);
    if (c->d_lma_f) (void)hipFree(c->d_lma_f);
    if (c->d_lp_ceil) (void)hipFree(c->d_lp_ceil);
    c->surr[0].release();
    c->surr[1].release();
    if (c->d_surr_list) (void)hipFree(c->d_surr_list);
    if (c->d_surr_fb) (void)hipFree(c->d_surr_fb);
    if (c->d_surr_now) (void)hipFree(c->d_surr_now);
    if (c->d_surr_perm) (void)hipFree(c->d_surr_perm);
    if (c->d_surr_bucket) (void)hipFree(c->d_surr_bucket);
    free_buf(c->defq);
    if (c->d_def_extra) (void)hipFree(c->d_def_extra);
    if (c->d_lasers) (void)hipFree(c->d_lasers);
    if (c->d_coll) (void)hipFree(c->d_coll);
    if (c->d_cdf) (void)hipFree(c->d_cdf);
    if (c->d_bw) (void)hipFree(c->d_bw);
    if (c->d_cp) (void)hipFree(c->d_cp);
    if (c->d_ev_extra) (void)hipFree(c->d_ev_extra);
    free_buf(c->evq); free_buf(c->lq[0]); free_buf(c->lq[1]); free_buf(c->phq); free_buf(c->phq2); free_buf(c->pairq);
    if (c->d_n0) (void)hipFree(c->d_n0);
    if (c->d_shower_count) (void)hipFree(c->d_shower_count);
    if (c->d_ctr) (void)hipFree(c->d_ctr);
    if (c->d_args) (void)hipFree(c->d_args);
    if (c->d_gargs) (void)hipFree(c->d_gargs);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->d_hist) (void)hipFree(c->d_hist);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* ptg_last_error(const ptg_ctx* c) { return c ? c->err.c_str() : "null context"; }

int ptg_set_laser(ptg_ctx* c, const ptg_laser_desc* d) {
    if (!c || !d) return fail(c, PTG_E_ARG, "null argument");
    LaserDev L;
    int rc = make_laser(c, d, L);
    if (rc) return rc;
    c->laser = L;
    c->have_laser = true;
    c->n_coll = 0;
    c->pol_angle = d->polarization == PTG_CIRCULAR ? 0.0 : d->pol_angle;
    c->lma_bandwidth = 0.0;
    if (d->finite_bandwidth) {  // plane_wave.rs:49-63, focused_laser.rs:45-58
        double nc = d->kind == PTG_FOCUSED_LASER ? L.n_cycles : d->n_cycles;
        double n_fwhm = d->envelope == PTG_COS_SQUARED ? 0.36405666377387671305 * nc : nc;
        c->lma_bandwidth = std::sqrt(0.5 * LN2_D) / (PI_H * n_fwhm);
    }
    return PTG_OK;
}

// On-axis energy density 0.5 eps0 (E^2 + c^2 B^2) of the FastPlaneWave at phase phi
// (fast_plane_wave.rs:110-185). As the reference's field, the amplitude carries the
// exact c^2 (SPEED_OF_LIGHT_SQD) and B = E / c with the 7-digit c, so c^2 B^2 differs
// from E^2 by (c_sqd / c^2).
static double fpw_axis_energy_density(const ptg_laser_desc* d, double k0, double phi) {
    const double amplitude = (M_ELECTRON * C_LIGHT_SQD * k0 * d->a0) / E_CHARGE;
    const double nc = d->n_cycles, b = d->chirp_b;
    const double psi = phi * (1.0 + b * phi), dpsi = 1.0 + 2.0 * b * phi;
    double f, df;
    if (d->envelope == PTG_COS_SQUARED) {
        if (std::fabs(phi) < nc * PI_H) {
            f = std::pow(std::cos(phi / (2.0 * nc)), 2);
            df = -std::sin(phi / nc) / (2.0 * nc);
        } else { f = 0.0; df = 0.0; }
    } else if (d->envelope == PTG_FLATTOP) {
        const double ap = std::fabs(phi);
        if (ap > PI_H * (nc + 1.0)) { f = 0.0; df = 0.0; }
        else if (ap > PI_H * (nc - 1.0)) {
            const double arg = 0.25 * (ap - (nc - 1.0) * PI_H);
            f = std::pow(std::cos(arg), 2);
            df = -std::copysign(1.0, phi) * 0.25 * std::sin(2.0 * arg);
        } else { f = 1.0; df = 0.0; }
    } else {
        const double q = phi / (PI_H * nc);
        f = std::exp2(-0.5 * q * q);
        df = -LN2_D * phi * f / ((PI_H * nc) * (PI_H * nc));
    }
    const double sn = std::sin(psi), cs = std::cos(psi);
    const double dax = sn * df + cs * dpsi * f;
    const double day = d->polarization == PTG_CIRCULAR ? (cs * df - sn * dpsi * f) : 0.0;
    const double e_sqd = amplitude * amplitude * (dax * dax + day * day);
    const double b_sqd = (amplitude / C_LIGHT) * (amplitude / C_LIGHT) * (dax * dax + day * day);
    return 0.5 * EPS0_SI * (e_sqd + C_LIGHT_SQD * b_sqd);
}

// FastPlaneWave::integrated_intensity (fast_plane_wave.rs:56-85): Simpson-weighted,
// compensated sum of the on-axis energy density over the pulse.
static double fpw_integrated_intensity(const ptg_laser_desc* d, double k0, int ppw) {
    const double nc = d->n_cycles;
    const double max_phase = d->envelope == PTG_COS_SQUARED ? PI_H * nc
                             : (d->envelope == PTG_FLATTOP ? PI_H * (nc + 1.0) : 6.0 * PI_H * nc);
    const double dphi = 2.0 * PI_H / double(ppw), dz = dphi / k0;
    const int64_t n_pts = int64_t(max_phase / dphi);
    double sum = 0.0, comp = 0.0;
    for (int64_t n = -n_pts; n < n_pts; ++n) {
        const double u = fpw_axis_energy_density(d, k0, double(n) * dphi);
        const double w = (n % 2 == 0) ? 2.0 / 3.0 : 4.0 / 3.0;
        const double y = w * u * dz - comp;
        const double t = sum + y;
        comp = (t - sum) - y;
        sum = t;
    }
    return sum;
}

int ptg_laser_energy(const ptg_laser_desc* d, double* energy, int32_t* per_area) {
    if (!d || !energy) return PTG_E_ARG;
    if (d->kind < 0 || d->kind > 3 || d->envelope < 0 || d->envelope > 2) return PTG_E_ARG;
    if (!(d->a0 > 0.0) || !(d->wavelength > 0.0) || !(d->n_cycles > 0.0)) return PTG_E_ARG;
    const bool focused = d->kind == PTG_FOCUSED_LASER || d->kind == PTG_FAST_FOCUSED_LASER;
    if (focused && !(d->waist > 0.0)) return PTG_E_ARG;
    const double k0 = 2.0 * PI_H / d->wavelength, omega = C_LIGHT * k0;
    const double amplitude = (M_ELECTRON * C_LIGHT * omega * d->a0) / E_CHARGE;
    // the focused fields' n_cycles is recomputed from their duration (fast_focused_laser.rs:47-53)
    const double nc = focused ? C_LIGHT * (d->n_cycles * d->wavelength / C_LIGHT) * k0 / (2.0 * PI_H) : d->n_cycles;
    if (per_area) *per_area = focused ? 0 : 1;
    // a chirped plane wave (the LMA plane wave keeps its chirp; the focused fields drop it)
    if (!focused && d->chirp_b != 0.0) {
        const int ppw = int(10.0 * (1.0 + 2.0 * PI_H * d->chirp_b * d->n_cycles));
        *energy = fpw_integrated_intensity(d, k0, ppw);
        return PTG_OK;
    }
    const double intensity = C_LIGHT * EPS0_SI * amplitude * amplitude;
    const double delta = d->polarization == PTG_CIRCULAR ? 1.0 : 0.0;
    double duration;
    if (d->envelope == PTG_COS_SQUARED) {
        duration = (1.0 + delta) * ((1.0 + 3.0 * nc * nc) * PI_H / (8.0 * nc)) / omega;
    } else if (d->envelope == PTG_GAUSSIAN) {
        const double arg = -(PI_H * nc) * (PI_H * nc) / LN2_D;
        const double large = 0.5 * nc * std::sqrt(PI_H * PI_H * PI_H / LN2_D);
        const double phase_x = large - std::expm1(arg) * std::sqrt(LN2_D / PI_H) / (4.0 * nc);
        const double phase_y = large + (1.0 + std::exp(arg)) * std::sqrt(LN2_D / PI_H) / (4.0 * nc);
        duration = (phase_x + delta * phase_y) / omega;
    } else {
        duration = (1.0 + delta) * ((nc - 3.0 / 16.0) * PI_H) / omega;
    }
    if (!focused) {
        *energy = intensity * duration;
        return PTG_OK;
    }
    const double zr = 0.5 * k0 * d->waist * d->waist, e = d->waist / zr;
    const double power = 0.5 * PI_H * intensity * d->waist * d->waist * (1.0 + e * e / 4.0 + std::pow(e, 4) / 8.0);
    *energy = power * duration;
    return PTG_OK;
}

int ptg_field_info(ptg_ctx* c, double* max_timestep, double* z0) {
    if (!c) return PTG_E_ARG;
    if (!c->have_laser) return fail(c, PTG_E_STATE, "no laser set");
    if (max_timestep) *max_timestep = c->laser.dt;
    if (z0) {
        ptg_laser_desc d{};
        d.envelope = c->laser.envelope;
        *z0 = ideal_initial_z(&d, c->laser);
    }
    return PTG_OK;
}

static void set_spectrum(const ptg_ctx* c, BeamDev& bd) {
    bd.spectrum = c->spectrum;
    bd.n_cdf = c->n_cdf;
    bd.cdf = c->d_cdf;
    bd.brem_min = c->brem_min;
    bd.brem_max = c->brem_max;
}

int ptg_set_beam_spectrum(ptg_ctx* c, int kind, const double* table, size_t n, double brem_min, double brem_max) {
    if (!c) return fail(c, PTG_E_ARG, "null argument");
    if (kind < 0 || kind > 2) return fail(c, PTG_E_ARG, "kind must be 0 (normal), 1 (tabulated) or 2 (bremsstrahlung)");
    if (kind == 1) {
        if (!table || n < 2 || n > (1u << 20)) return fail(c, PTG_E_ARG, "a tabulated spectrum needs 2 to 2^20 points");
        for (size_t i = 0; i < n; ++i) {
            if (!std::isfinite(table[2 * i]) || !std::isfinite(table[2 * i + 1]))
                return fail(c, PTG_E_ARG, "the spectrum holds non-numerical values");
            if (i && (table[2 * i] <= table[2 * i - 2] || table[2 * i + 1] < table[2 * i - 1]))
                return fail(c, PTG_E_ARG, "the spectrum's gamma must increase and its CDF must not decrease");
        }
        HIPCHK(c, hipSetDevice(c->device));
        if (c->d_cdf) (void)hipFree(c->d_cdf);
        c->d_cdf = nullptr;
        HIPCHK(c, hipMalloc(&c->d_cdf, n * 2 * sizeof(double)));
        HIPCHK(c, scopy(c, c->d_cdf, table, n * 2 * sizeof(double), hipMemcpyHostToDevice));
        c->n_cdf = int(n);
    }
    if (kind == 2 && !(brem_min > 0.0 && brem_max > brem_min))
        return fail(c, PTG_E_ARG, "a bremsstrahlung source needs 0 < gamma_min < gamma");
    c->spectrum = kind;
    c->brem_min = brem_min;
    c->brem_max = brem_max;
    return PTG_OK;
}

int ptg_build_beam(ptg_ctx* c, const ptg_beam_desc* b, double initial_z, uint64_t seed) {
    if (!c || !b) return fail(c, PTG_E_ARG, "null argument");
    if (b->n == 0 || b->n > 0xffffffffull) return fail(c, PTG_E_ARG, "beam size must be in [1, 2^32)");
    if (b->species < 0 || b->species > 2 || b->radial < 0 || b->radial > 2) return fail(c, PTG_E_ARG, "bad species or radial distribution");
    if (!(std::fabs(b->energy_chirp) <= 1.0)) return fail(c, PTG_E_ARG, "|energy_chirp| must be <= 1");
    HIPCHK(c, hipSetDevice(c->device));
    if (c->prim.cap < b->n) HIPCHK(c, alloc_buf(c->prim, b->n));
    int rc = ensure_n0(c, b->n);
    if (rc) return rc;
    BeamDev bd{};
    set_spectrum(c, bd);
    bd.n = uint32_t(b->n);
    bd.species = b->species;
    bd.radial = b->radial;
    bd.gamma = b->gamma;
    bd.sigma = b->sigma;
    bd.rho = -b->energy_chirp;
    bd.radius = b->radius;
    bd.max_radius = b->max_radius;
    bd.length = b->length;
    bd.rms_div = b->rms_div;
    bd.angle = b->angle;
    bd.plane_angle = b->collision_plane_angle;
    bd.off_x = b->offset[0]; bd.off_y = b->offset[1]; bd.off_z = b->offset[2];
    for (int k = 0; k < 4; ++k) bd.pol[k] = b->pol[k];
    bd.weight = b->weight;
    bd.initial_z = initial_z;
    bd.key = mix64(key_from_seed(seed) ^ 0x6265616d6265616dULL);
    bd.off = 0;
    bd.n_total = bd.n;
    c->n_coll = 0;
    launch_build_beam(bd, c->prim.v, c->d_n0, c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->n_prim = b->n;
    return PTG_OK;
}

static void soa_from_host(std::vector<double>* cols, std::vector<uint64_t>* u64, std::vector<uint32_t>* u32,
                          const ptg_particles* p, size_t n, bool primaries) {
    for (int k = 0; k < P_NDOUBLE; ++k) cols[k].assign(n, 0.0);
    for (int k = 0; k < 3; ++k) u64[k].assign(n, 0);
    for (int k = 0; k < 3; ++k) u32[k].assign(n, 0);
    for (size_t i = 0; i < n; ++i) {
        for (int c4 = 0; c4 < 4; ++c4) {
            cols[P_RT + c4][i] = p->r[4 * i + c4];
            cols[P_UT + c4][i] = p->u[4 * i + c4];
            cols[P_R0T + c4][i] = p->r0 ? p->r0[4 * i + c4] : p->r[4 * i + c4];
            double u0def = (c4 == 0 && p->species[i] != 2) ? 1.0 : 0.0;
            cols[P_U0T + c4][i] = p->u0 ? p->u0[4 * i + c4] : u0def;
            cols[P_SI + c4][i] = p->pol ? p->pol[4 * i + c4] : (c4 == 0 ? 1.0 : 0.0);
        }
        cols[P_PAYLOAD][i] = p->payload ? p->payload[i] : 0.0;
        cols[P_PCHI][i] = p->parent_chi ? p->parent_chi[i] : 0.0;
        cols[P_ICOUNT][i] = p->interaction_count ? p->interaction_count[i] : 0.0;
        cols[P_WORK][i] = p->work ? p->work[i] : 0.0;
        cols[P_WEIGHT][i] = p->weight ? p->weight[i] : 1.0;
        u64[0][i] = p->id ? p->id[i] : i;
        u64[1][i] = p->parent_id ? p->parent_id[i] : u64[0][i];
        u64[2][i] = i;  // RNG stream of primary i
        u32[0][i] = uint32_t(i);
        u32[1][i] = uint32_t(p->species[i]) | (primaries ? META_PRIMARY : 0u);
        u32[2][i] = 0;
    }
}

int ptg_load_primaries(ptg_ctx* c, const ptg_particles* p, size_t n) {
    if (!c || !p) return fail(c, PTG_E_ARG, "null argument");
    if (n == 0 || n > 0xffffffffull) return fail(c, PTG_E_ARG, "number of primaries must be in [1, 2^32)");
    if (!p->species || !p->r || !p->u) return fail(c, PTG_E_ARG, "species, r and u are required");
    for (size_t i = 0; i < n; ++i)
        if (p->species[i] > 2) return fail(c, PTG_E_ARG, "species out of range");
    HIPCHK(c, hipSetDevice(c->device));
    if (c->prim.cap < n) HIPCHK(c, alloc_buf(c->prim, n));
    int rc = ensure_n0(c, n);
    if (rc) return rc;
    std::vector<double> cols[P_NDOUBLE];
    std::vector<uint64_t> u64[3];
    std::vector<uint32_t> u32[3];
    soa_from_host(cols, u64, u32, p, n, true);
    for (int k = 0; k < P_NDOUBLE; ++k) HIPCHK(c, scopy(c, c->prim.v.d[k], cols[k].data(), n * 8, hipMemcpyHostToDevice));
    HIPCHK(c, scopy(c, c->prim.v.id, u64[0].data(), n * 8, hipMemcpyHostToDevice));
    HIPCHK(c, scopy(c, c->prim.v.parent_id, u64[1].data(), n * 8, hipMemcpyHostToDevice));
    HIPCHK(c, scopy(c, c->prim.v.stream, u64[2].data(), n * 8, hipMemcpyHostToDevice));
    HIPCHK(c, scopy(c, c->prim.v.shower, u32[0].data(), n * 4, hipMemcpyHostToDevice));
    HIPCHK(c, scopy(c, c->prim.v.meta, u32[1].data(), n * 4, hipMemcpyHostToDevice));
    HIPCHK(c, scopy(c, c->prim.v.nev, u32[2].data(), n * 4, hipMemcpyHostToDevice));
    launch_primary_dirs(c->prim.v, uint32_t(n), c->d_n0, c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->n_coll = 0;
    c->n_prim = n;
    return PTG_OK;
}

static int validate_options(ptg_ctx* c, const ptg_collide_options* o) {
    if (!(o->dt_multiplier > 0.0)) return fail(c, PTG_E_ARG, "dt_multiplier must be positive");
    if (!(o->rate_increase >= 1.0)) return fail(c, PTG_E_ARG, "increase in pair creation rate must be >= 1.0");
    if (o->rate_increase > 1.0 && o->multiplicity >= 0)
        return fail(c, PTG_E_ARG, "event biasing (pair rate increase > 1) is incompatible with multiplicity selection");
    if (o->gaunt_factor && !o->classical) return fail(c, PTG_E_ARG, "gaunt_factor requires classical");
    if (o->precision != PTG_PREC_FP64 && o->precision != PTG_PREC_MIXED) return fail(c, PTG_E_ARG, "unknown precision mode");
    return PTG_OK;
}

// One transport launch over `n_in` particles, repeated with a larger secondary
// queue if the first attempt overflowed it (the launch is deterministic, so the
// repeat reproduces the same particles).
static int transport(ptg_ctx* c, TransportArgs a, bool photons, DevBuf& in, size_t n_in, DevBuf& next, double* ms,
                     uint64_t* launches) {
    if (n_in == 0) return PTG_OK;
    // room for one final-state record per input particle
    HIPCHK(c, grow_buf(c->out, c->n_out + n_in + n_in / 4 + 1024, c->n_out, c->stream));
    if (a.opt.keep_decayed_photons && photons) HIPCHK(c, grow_buf(c->inter, c->n_inter + n_in + 1024, c->n_inter, c->stream));
    {
        const double ratio = c->next_ratio[photons ? 1 : 0];
        const size_t want = size_t(double(n_in) * ratio * 1.25) + 1024;
        if (next.cap < want) HIPCHK(c, alloc_buf(next, want));
    }
    unsigned long long snap[C_N];
    HIPCHK(c, scopy(c, snap, c->d_ctr, sizeof(snap), hipMemcpyDeviceToHost));
    for (int attempt = 0; attempt < 4; ++attempt) {
        snap[C_WORK] = 0;
        snap[C_NEXT] = 0;
        HIPCHK(c, hipMemcpyAsync(c->d_ctr, snap, sizeof(snap), hipMemcpyHostToDevice, c->stream));
        a.in = in.v;
        a.n_in = uint32_t(n_in);
        a.out = c->out.v;
        a.inter = c->inter.v;
        a.next = next.v;
        a.next_cap = next.cap;
        int blocks_per_cu = lcfa_blocks_per_cu(a.laser.kind, a.laser.circular, photons);
        if (const char* e = std::getenv("PTG_BLOCKS_PER_CU")) {  // diagnostic override for grid sweeps
            int v = std::atoi(e);
            if (v >= 1 && v <= 16) blocks_per_cu = v;
        }
        int grid = int(std::min<size_t>(size_t(c->n_cu) * blocks_per_cu, (n_in + 255) / 256));
        HIPCHK(c, hipEventRecord(c->ev0, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->d_args, &a, sizeof(a), hipMemcpyHostToDevice, c->stream));
        if (photons) launch_photons_lcfa(c->d_args, a.laser.kind, a.laser.circular, grid, c->stream);
        else launch_leptons_lcfa(c->d_args, a.laser.kind, a.laser.circular, grid, c->stream);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
        unsigned long long now[C_N];
        HIPCHK(c, hipMemcpyAsync(now, c->d_ctr, sizeof(now), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        float t = 0.0f;
        (void)hipEventElapsedTime(&t, c->ev0, c->ev1);
        *ms += t;
        ++*launches;
        const double seen = double(now[C_NEXT]) / double(n_in);
        double& ratio = c->next_ratio[photons ? 1 : 0];
        if (seen > ratio) ratio = seen;
        if (now[C_NEXT] <= next.cap) {
            c->n_out = now[C_OUT];
            c->n_inter = now[C_INTER];
            return PTG_OK;
        }
        // overflow: grow the queue and repeat from the snapshot
        HIPCHK(c, alloc_buf(next, size_t(now[C_NEXT] + now[C_NEXT] / 8 + 1024)));
    }
    return fail(c, PTG_E_OOM, "secondary queue kept overflowing");
}

static int ensure_ev_extra(ptg_ctx* c, size_t n) {
    if (n <= c->ev_extra_cap) return PTG_OK;
    if (c->d_ev_extra) (void)hipFree(c->d_ev_extra);
    c->d_ev_extra = nullptr;
    HIPCHK(c, hipMalloc(&c->d_ev_extra, 2 * n * sizeof(double)));
    c->ev_extra_cap = n;
    return PTG_OK;
}

// Runs the LMA event kernel (emissions, or pairs when `pairs`) over the n_ev
// events in evq (or `list` of `in`).
static int lma_event_pass(ptg_ctx* c, LmaArgs m, bool pairs, size_t n_ev, size_t max_lanes, double* ms, uint64_t* launches,
                          const uint32_t* list = nullptr, const DevBuf* in = nullptr) {
    m.ev_in = in ? in->v : c->evq.v;
    m.ev_list = list;
    m.n_ev = n_ev;
    // one event per wave
    const size_t lanes = std::min<size_t>(max_lanes, n_ev * 64);
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    if (pairs) launch_pair_event_lma(c->d_args, m, int(lanes), c->stream);
    else launch_event_lma(c->d_args, m, int(lanes), c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float t = 0.0f;
    (void)hipEventElapsedTime(&t, c->ev0, c->ev1);
    *ms += t;
    ++*launches;
    return PTG_OK;
}

// Box of the surrogate's domain in SurrArgs terms
static void surr_box_args(const double* box, SurrArgs& S) {
    const double la0 = std::log(box[0]), la1 = std::log(box[1]), le0 = std::log(box[2]), le1 = std::log(box[3]);
    S.la_c = 0.5 * (la0 + la1);
    S.la_h = 0.5 * (la1 - la0);
    S.le_c = 0.5 * (le0 + le1);
    S.le_h = 0.5 * (le1 - le0);
    S.la_use = std::log(std::max(box[0], box[4]));
    S.le_use = std::log(std::max(box[2], box[5]));
}

// Index lists of the learned sampler's pass (rows, deferred, exact now, order)
// for n_ev events, and its bucket counts.
static int ensure_surr_lists(ptg_ctx* c, size_t n_ev) {
    if (c->surr_list_cap < n_ev) {
        if (c->d_surr_list) (void)hipFree(c->d_surr_list);
        if (c->d_surr_fb) (void)hipFree(c->d_surr_fb);
        if (c->d_surr_now) (void)hipFree(c->d_surr_now);
        if (c->d_surr_perm) (void)hipFree(c->d_surr_perm);
        c->d_surr_list = c->d_surr_fb = c->d_surr_now = c->d_surr_perm = nullptr;
        c->surr_list_cap = 0;
        const size_t cap = n_ev + n_ev / 4 + 64;
        HIPCHK(c, hipMalloc(&c->d_surr_list, cap * sizeof(uint32_t)));
        HIPCHK(c, hipMalloc(&c->d_surr_fb, cap * sizeof(uint32_t)));
        HIPCHK(c, hipMalloc(&c->d_surr_now, cap * sizeof(uint32_t)));
        HIPCHK(c, hipMalloc(&c->d_surr_perm, cap * sizeof(uint32_t)));
        c->surr_list_cap = cap;
    }
    if (!c->d_surr_bucket) HIPCHK(c, hipMalloc(&c->d_surr_bucket, 64 * sizeof(unsigned int)));
    return PTG_OK;
}

// Room for `need` events in the deferred queue (its records, kappa^0 and
// sqrt(<a^2>)), keeping the n_def held.
static int ensure_deferred(ptg_ctx* c, size_t need) {
    if (c->defq.cap < need) HIPCHK(c, grow_buf(c->defq, need + need / 2 + 64, c->n_def, c->stream));
    if (c->def_extra_cap < c->defq.cap) {
        double* ne = nullptr;
        HIPCHK(c, hipMalloc(&ne, 2 * c->defq.cap * sizeof(double)));
        if (c->n_def) {
            HIPCHK(c, hipMemcpyAsync(ne, c->d_def_extra, c->n_def * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(ne + c->defq.cap, c->d_def_extra + c->def_extra_cap, c->n_def * sizeof(double),
                                     hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        }
        if (c->d_def_extra) (void)hipFree(c->d_def_extra);
        c->d_def_extra = ne;
        c->def_extra_cap = c->defq.cap;
    }
    return PTG_OK;
}

// Emission events of one generation through the learned sampler: inputs for
// the events inside its domain, the network's forward pass, the photons; the
// other events (outside the domain, or an LP harmonic beyond the workspace) go
// to the exact kernels.
static int lma_surrogate_pass(ptg_ctx* c, LmaArgs m, size_t n_ev, size_t max_lanes, double* ms, uint64_t* launches,
                              uint64_t* n_fallback) {
    const int p = m.mode == LMA_CP ? 1 : 0;
    MlpNet& net = c->surr[p];
    int rc = ensure_surr_lists(c, n_ev);
    if (rc) return rc;
    const size_t rows_cap = (n_ev + 31) / 32 * 32;
    if (rows_cap > size_t(1) << 30) return fail(c, PTG_E_OOM, "too many emission events for the surrogate's buffers");
    HIPCHK(c, net.reserve(int(rows_cap)));
    HIPCHK(c, hipMemsetAsync(c->d_ctr + C_SROWS, 0, 3 * sizeof(unsigned long long), c->stream));
    SurrArgs S{};
    surr_box_args(c->surr_box[p], S);
    S.x = net.X;
    S.out = net.O;
    S.list = c->d_surr_list;
    S.n_rows = c->d_ctr + C_SROWS;
    S.fb = c->d_surr_fb;
    S.n_fb = c->d_ctr + C_SFB;
    S.now = c->d_surr_now;
    S.n_now = c->d_ctr + C_SNOW;
    S.n_ev = n_ev;
    S.perm = c->d_surr_perm;
    S.bucket = c->d_surr_bucket;
    m.ev_in = c->evq.v;
    m.ev_list = nullptr;
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    launch_surr_inputs(c->d_args, m, S, c->stream);
    HIPCHK(c, hipGetLastError());
    unsigned long long cnt[3] = {0, 0, 0};
    HIPCHK(c, hipMemcpyAsync(cnt, c->d_ctr + C_SROWS, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const size_t n_rows = size_t(cnt[0]), n_now = size_t(cnt[2]);
    if (n_rows > 0) {
        const size_t M = (n_rows + 31) / 32 * 32;
        if (M > n_rows) HIPCHK(c, hipMemsetAsync(net.X + 8 * n_rows, 0, (M - n_rows) * 8 * sizeof(float), c->stream));
        net.forward(int(M), false, c->stream);
        HIPCHK(c, hipMemsetAsync(c->d_surr_bucket, 0, 64 * sizeof(unsigned int), c->stream));
        launch_surr_order(S, uint32_t(n_rows), c->stream);
        const size_t lanes = std::min<size_t>(max_lanes, (n_rows + 63) / 64 * 64);
        launch_surr_event(c->d_args, m, S, int(lanes), c->stream);
        HIPCHK(c, hipGetLastError());
    }
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipMemcpyAsync(cnt, c->d_ctr + C_SROWS, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float t = 0.0f;
    (void)hipEventElapsedTime(&t, c->ev0, c->ev1);
    *ms += t;
    ++*launches;
    const size_t n_fb = size_t(cnt[1]);
    *n_fallback += n_fb + n_now;
    if (n_now > 0) {  // below the domain's amplitude: the exact kernels, in this generation
        rc = lma_event_pass(c, m, false, n_now, max_lanes, ms, launches, c->d_surr_now, &c->evq);
        if (rc) return rc;
    }
    if (n_fb > 0) {  // held for one exact pass when the lepton phase drains
        rc = ensure_deferred(c, c->n_def + n_fb);
        if (rc) return rc;
        launch_gather_events(c->evq.v, m.ev_k0, m.ev_a, c->d_surr_fb, n_fb, c->defq.v, c->d_def_extra,
                             c->d_def_extra + c->def_extra_cap, c->n_def, c->stream);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipStreamSynchronize(c->stream));
        c->n_def += n_fb;
    }
    return PTG_OK;
}

static unsigned long long read_ctr(ptg_ctx* c, int which) {
    unsigned long long v = 0;
    if (scopy(c, &v, c->d_ctr + which, sizeof(v), hipMemcpyDeviceToHost) != hipSuccess) return ~0ull;
    return v;
}

// ---------------------------------------------------------------- captured generations
// PTG_LAUNCH_GRAPH: after its first generation, each phase of the LMA loop runs
// as sequences of generations captured into a HIP graph and launched at once.
// Every launch in a sequence reads its counts from device memory (the start of
// each generation, k_gen_begin, moves the count the last one queued into the
// launch arguments), so the host reads the counter block once per sequence
// instead of after every launch. A generation that could overrun an output
// queue stalls the rest of its sequence; the host grows the queues and resumes.

// Generations per captured sequence (even: the queues alternate between two
// buffers, so that every sequence ends on the buffer it started from).
constexpr int GRAPH_GENS = 16;

// Diagnostic override (tests of the stall path): PTG_GRAPH_QUEUE_ROOM = records
// of room a sequence may use in each output queue beyond what it holds, however
// large the queue; 0 or unset: the queues' capacities.
static size_t graph_room_override() {
    const char* e = std::getenv("PTG_GRAPH_QUEUE_ROOM");
    if (!e) return 0;
    const long long v = std::atoll(e);
    return v > 0 ? size_t(v) : 0;
}
static unsigned long long room_cap(size_t cap, size_t held, size_t over) {
    return over ? std::min<unsigned long long>(cap, held + over) : cap;
}

// Captures what `record` launches on the context stream, makes it executable
// and launches it between ev0 and ev1.
static int graph_launch(ptg_ctx* c, const std::function<void()>& record) {
    hipGraph_t g = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    record();
    const hipError_t le = hipGetLastError();
    const hipError_t ee = hipStreamEndCapture(c->stream, &g);
    if (le != hipSuccess || ee != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(c, PTG_E_HIP, std::string("graph capture: ") + hipGetErrorString(le != hipSuccess ? le : ee));
    }
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    const hipError_t ie = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
        c->gexec = nullptr;
        return fail(c, PTG_E_HIP, std::string("graph instantiation: ") + hipGetErrorString(ie));
    }
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    ++c->graph_launches;
    return PTG_OK;
}

// Reads the counter block after a sequence; returns the device time.
static int graph_finish(ptg_ctx* c, unsigned long long* now, double* ms) {
    HIPCHK(c, hipMemcpyAsync(now, c->d_ctr, C_N * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float t = 0.0f;
    (void)hipEventElapsedTime(&t, c->ev0, c->ev1);
    *ms += t;
    return PTG_OK;
}

// Sets the counters a sequence starts from: its input count (C_LNEXT), the
// deferred events held (C_NDEF), no stall, no generation run yet (C_GENS).
static int graph_prime(ptg_ctx* c, size_t n_in) {
    unsigned long long v[2] = {n_in, 0ull};
    HIPCHK(c, hipMemcpyAsync(c->d_ctr + C_LNEXT, &v[0], sizeof(v[0]), hipMemcpyHostToDevice, c->stream));
    unsigned long long nd[1] = {c->n_def};
    HIPCHK(c, hipMemcpyAsync(c->d_ctr + C_NDEF, nd, sizeof(nd), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_ctr + C_STALL, &v[1], sizeof(v[1]), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_ctr + C_GENS, &v[1], sizeof(v[1]), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PTG_OK;
}

// Lepton generations of the LMA loop from a captured sequence, the input in
// lq[s] (n_cur electrons and positrons). On return *cur and *n_left hold what
// is left and *li the buffer the next generation writes. A generation whose
// events found too little room in the photon or final-state queues (the
// sequence stalled there) has its events sampled here, eagerly, after the
// queues have grown.
static int lepton_sequence(ptg_ctx* c, TransportArgs a, LmaArgs m, bool track, bool use_surr, int s, size_t n_cur,
                           size_t max_lanes, double* ms, uint64_t* launches, uint64_t* gens, uint64_t* surr_fb, DevBuf** cur,
                           size_t* n_left, int* li) {
    // every lepton of the phase leaves the field once (a final state); each event
    // adds a photon (queued, or a final state) and re-queues its lepton. Room for
    // the photons of four generations in which every lepton emits; the device
    // checks the room before each generation's events
    const size_t budget = 4 * n_cur + 4096;
    HIPCHK(c, grow_buf(c->out, c->n_out + n_cur + (track ? 0 : budget) + 1024, c->n_out, c->stream));
    if (c->evq.cap < n_cur) HIPCHK(c, alloc_buf(c->evq, n_cur + n_cur / 4));
    int rc = ensure_ev_extra(c, c->evq.cap);
    if (rc) return rc;
    if (c->lq[s ^ 1].cap < n_cur) HIPCHK(c, alloc_buf(c->lq[s ^ 1], n_cur + n_cur / 4));
    const unsigned long long n_ph = read_ctr(c, track ? C_PH : C_NEXT);
    HIPCHK(c, grow_buf(c->phq, size_t(n_ph) + (track ? budget : 0) + 1024, size_t(n_ph), c->stream));

    const int p = m.mode == LMA_CP ? 1 : 0;
    MlpNet& net = c->surr[p];
    SurrArgs S{};
    const size_t rows_cap = (n_cur + 31) / 32 * 32;
    if (use_surr) {
        if (rows_cap > size_t(1) << 30) return fail(c, PTG_E_OOM, "too many emission events for the surrogate's buffers");
        rc = ensure_surr_lists(c, n_cur);
        if (rc) return rc;
        HIPCHK(c, net.reserve(int(rows_cap)));
        rc = ensure_deferred(c, c->n_def + n_cur);
        if (rc) return rc;
        surr_box_args(c->surr_box[p], S);
        S.x = net.X;
        S.out = net.O;
        S.list = c->d_surr_list;
        S.n_rows = c->d_ctr + C_SROWS;
        S.fb = c->d_surr_fb;
        S.n_fb = c->d_ctr + C_SFB;
        S.now = c->d_surr_now;
        S.n_now = c->d_ctr + C_SNOW;
        S.n_ev = n_cur;
        S.n_ev_dev = c->d_ctr + C_EVRUN;
        S.perm = c->d_surr_perm;
        S.bucket = c->d_surr_bucket;
    }

    // launch arguments of the two parities: slot k reads lq[k] and writes lq[k ^ 1]
    TransportArgs ta[2];
    LmaArgs mk[2];
    for (int k = 0; k < 2; ++k) {
        ta[k] = a;
        ta[k].in = c->lq[k].v;
        ta[k].n_in = 0;
        ta[k].out = c->out.v;
        ta[k].inter = c->inter.v;
        ta[k].next = c->phq.v;
        ta[k].next_ctr = c->d_ctr + (track ? C_PH : C_NEXT);
        ta[k].next_cap = c->phq.cap;
        mk[k] = m;
        mk[k].ev = c->evq.v;
        mk[k].ev_cap = c->evq.cap;
        mk[k].ev_k0 = c->d_ev_extra;
        mk[k].ev_a = c->d_ev_extra + c->ev_extra_cap;
        mk[k].ph = c->phq.v;
        mk[k].ph_cap = c->phq.cap;
        mk[k].ev_in = c->evq.v;
        mk[k].ev_list = nullptr;
        mk[k].n_ev = 0;
        mk[k].n_ev_dev = c->d_ctr + C_EVRUN;
        mk[k].next = c->lq[k ^ 1].v;
        mk[k].next_ctr = c->d_ctr + C_LNEXT;
    }
    HIPCHK(c, scopy(c, c->d_gargs, ta, sizeof(ta), hipMemcpyHostToDevice));
    // per event: the photon when it is a final state, and room kept for the final
    // state of the lepton it re-queues (the step kernel writes final states unchecked)
    const size_t over = graph_room_override();
    EventRoom room{};
    room.out_cap = room_cap(c->out.cap, c->n_out + n_cur, over);
    room.out_per = track ? 1u : 2u;
    room.ph_cap = room_cap(c->phq.cap, size_t(n_ph), over);
    room.ph_per = track ? 1u : 0u;
    room.ph_slot = track ? C_PH : C_NEXT;
    room.pairs_cap = ~0ull;
    room.inter_cap = ~0ull;
    rc = graph_prime(c, n_cur);
    if (rc) return rc;

    const size_t sb = size_t(c->step_block);
    const int grid = int(std::min<size_t>(size_t(c->n_cu) * size_t(c->step_grid_per_cu), (n_cur + sb - 1) / sb));
    const int ev_lanes = int(std::min<size_t>(max_lanes, n_cur * 64));
    const int surr_lanes = int(std::min<size_t>(max_lanes, (n_cur + 63) / 64 * 64));
    rc = graph_launch(c, [&]() {
        for (int g = 0; g < GRAPH_GENS; ++g) {
            const int k = s ^ (g & 1);
            TransportArgs* A = c->d_gargs + k;
            launch_gen_begin(c->d_ctr, A, use_surr ? c->d_surr_bucket : nullptr, c->stream);
            launch_step_lma(A, mk[k], grid, c->stream, c->step_block);
            launch_event_room(c->d_ctr, room, c->stream);
            if (!use_surr) {
                launch_event_lma(A, mk[k], ev_lanes, c->stream);
                continue;
            }
            // the learned sampler: network rows, forward pass, rows ordered by
            // harmonic, photons; the deferred events appended to the deferred
            // queue (those without room, and the events below the network's
            // range, to the exact kernels in this generation)
            launch_surr_inputs(A, mk[k], S, c->stream);
            launch_surr_pad(S, c->stream);
            net.forward(int(rows_cap), false, c->stream, S.n_rows);
            launch_surr_order_dev(S, uint32_t(n_cur), c->stream);
            launch_surr_event(A, mk[k], S, surr_lanes, c->stream);
            launch_gather_deferred(c->evq.v, c->d_ev_extra, c->d_ev_extra + c->ev_extra_cap, c->d_surr_fb, S.n_fb, n_cur,
                                   c->defq.v, c->d_def_extra, c->d_def_extra + c->def_extra_cap, c->d_ctr + C_NDEF,
                                   c->defq.cap, c->d_surr_now, S.n_now, c->stream);
            launch_commit_deferred(S.n_fb, c->d_ctr + C_NDEF, c->defq.cap, S.n_now, c->d_ctr + C_STATS, c->stream);
            LmaArgs mn = mk[k];
            mn.ev_list = c->d_surr_now;
            mn.n_ev_dev = S.n_now;
            launch_event_lma(A, mn, ev_lanes, c->stream);
        }
    });
    if (rc) return rc;
    unsigned long long now[C_N];
    rc = graph_finish(c, now, &ms[0]);
    if (rc) return rc;
    ++launches[0];
    c->n_out = now[C_OUT];
    c->n_def = size_t(now[C_NDEF]);
    // generations that ran their step kernel; the last one's output is the input left
    const uint64_t ran = now[C_GENS];
    *gens += ran;
    int last = int((uint64_t(s) + ran) & 1u);
    if (now[C_STALL]) {
        // the last generation's events are still queued: grow the queues and sample them
        ++c->graph_stalls;
        const size_t n_ev = size_t(now[C_EV]);
        HIPCHK(c, grow_buf(c->out, c->n_out + (track ? 0 : n_ev) + 1024, c->n_out, c->stream));
        HIPCHK(c, grow_buf(c->phq, size_t(now[track ? C_PH : C_NEXT]) + (track ? n_ev : 0) + 1024,
                           size_t(now[track ? C_PH : C_NEXT]), c->stream));
        LmaArgs me = mk[last ^ 1];
        me.ph = c->phq.v;
        me.ph_cap = c->phq.cap;
        me.n_ev_dev = nullptr;
        TransportArgs te = ta[last ^ 1];
        te.out = c->out.v;
        te.next = c->phq.v;
        te.next_cap = c->phq.cap;
        HIPCHK(c, scopy(c, c->d_args, &te, sizeof(te), hipMemcpyHostToDevice));
        rc = use_surr ? lma_surrogate_pass(c, me, n_ev, max_lanes, &ms[1], &launches[1], surr_fb)
                      : lma_event_pass(c, me, false, n_ev, max_lanes, &ms[1], &launches[1]);
        if (rc) return rc;
        now[C_LNEXT] = read_ctr(c, C_LNEXT);
        c->n_out = read_ctr(c, C_OUT);
    }
    *cur = &c->lq[last];
    *li = last ^ 1;
    *n_left = size_t(now[C_LNEXT]);
    return PTG_OK;
}

// Photon generations of the LMA loop from a captured sequence, the input in
// phq (s = 0) or phq2 (s = 1), n_ph photons. On return *pcur and *n_left hold
// what is left; as lepton_sequence, a stalled generation's pair events are
// sampled here after the queues have grown.
static int photon_sequence(ptg_ctx* c, TransportArgs a, LmaArgs m, int s, size_t n_ph, size_t max_lanes, double* ms,
                           uint64_t* launches, uint64_t* gens, DevBuf** pcur, size_t* n_left) {
    DevBuf* pb[2] = {&c->phq, &c->phq2};
    const bool keep = a.opt.keep_decayed_photons != 0;
    // every photon of the phase leaves the field once (a final state) or decays;
    // each pair event writes two leptons and a decayed photon or re-queues the
    // photon. Room for a pair event per eight photons; the device checks the room
    // before each generation's events
    const size_t budget = n_ph / 8 + 4096;
    HIPCHK(c, grow_buf(c->out, c->n_out + n_ph + 1024, c->n_out, c->stream));
    if (keep) HIPCHK(c, grow_buf(c->inter, c->n_inter + budget, c->n_inter, c->stream));
    if (c->evq.cap < n_ph) HIPCHK(c, alloc_buf(c->evq, n_ph + n_ph / 4));
    int rc = ensure_ev_extra(c, c->evq.cap);
    if (rc) return rc;
    if (pb[s ^ 1]->cap < n_ph) HIPCHK(c, alloc_buf(*pb[s ^ 1], n_ph + n_ph / 4));
    const unsigned long long n_pairs = read_ctr(c, C_PAIRS);
    HIPCHK(c, grow_buf(c->pairq, size_t(n_pairs) + 2 * budget, size_t(n_pairs), c->stream));

    TransportArgs ta[2];
    LmaArgs mk[2];
    for (int k = 0; k < 2; ++k) {
        ta[k] = a;
        ta[k].in = pb[k]->v;
        ta[k].n_in = 0;
        ta[k].out = c->out.v;
        ta[k].inter = c->inter.v;
        ta[k].next = c->pairq.v;  // (no leptons are queued with photons)
        ta[k].next_ctr = c->d_ctr + C_PAIRS;
        ta[k].next_cap = c->pairq.cap;
        mk[k] = m;
        mk[k].ev = c->evq.v;
        mk[k].ev_cap = c->evq.cap;
        mk[k].ev_k0 = c->d_ev_extra;
        mk[k].ev_a = c->d_ev_extra + c->ev_extra_cap;
        mk[k].ev_in = c->evq.v;
        mk[k].ev_list = nullptr;
        mk[k].n_ev = 0;
        mk[k].n_ev_dev = c->d_ctr + C_EVRUN;
        mk[k].next = pb[k ^ 1]->v;
        mk[k].next_ctr = c->d_ctr + C_LNEXT;
        mk[k].pairs = c->pairq.v;
        mk[k].pairs_cap = c->pairq.cap;
    }
    HIPCHK(c, scopy(c, c->d_gargs, ta, sizeof(ta), hipMemcpyHostToDevice));
    // per event: two pair records, the decayed photon when kept, and room kept
    // for the final state of the photon when it is re-queued
    const size_t over = graph_room_override();
    EventRoom room{};
    room.out_cap = room_cap(c->out.cap, c->n_out + n_ph, over);
    room.out_per = 1u;
    room.ph_cap = ~0ull;
    room.ph_slot = C_PH;
    room.pairs_cap = room_cap(c->pairq.cap, size_t(n_pairs), over);
    room.pairs_per = 2u;
    room.inter_cap = keep ? room_cap(c->inter.cap, c->n_inter, over) : ~0ull;
    room.inter_per = keep ? 1u : 0u;
    rc = graph_prime(c, n_ph);
    if (rc) return rc;
    const int grid = int(std::min<size_t>(size_t(c->n_cu) * 4, (n_ph + 255) / 256));
    const int ev_lanes = int(std::min<size_t>(max_lanes, n_ph * 64));
    rc = graph_launch(c, [&]() {
        for (int g = 0; g < GRAPH_GENS; ++g) {
            const int k = s ^ (g & 1);
            TransportArgs* A = c->d_gargs + k;
            launch_gen_begin(c->d_ctr, A, nullptr, c->stream);
            launch_photon_step_lma(A, mk[k], grid, c->stream);
            launch_event_room(c->d_ctr, room, c->stream);
            launch_pair_event_lma(A, mk[k], ev_lanes, c->stream);
        }
    });
    if (rc) return rc;
    unsigned long long now[C_N];
    rc = graph_finish(c, now, &ms[1]);
    if (rc) return rc;
    ++launches[1];
    c->n_out = now[C_OUT];
    c->n_inter = now[C_INTER];
    const uint64_t ran = now[C_GENS];
    *gens += ran;
    const int last = int((uint64_t(s) + ran) & 1u);
    if (now[C_STALL]) {
        ++c->graph_stalls;
        const size_t n_ev = size_t(now[C_EV]);
        if (keep) HIPCHK(c, grow_buf(c->inter, c->n_inter + n_ev + 1024, c->n_inter, c->stream));
        HIPCHK(c, grow_buf(c->pairq, size_t(now[C_PAIRS]) + 2 * n_ev + 1024, size_t(now[C_PAIRS]), c->stream));
        LmaArgs me = mk[last ^ 1];
        me.pairs = c->pairq.v;
        me.pairs_cap = c->pairq.cap;
        me.n_ev_dev = nullptr;
        TransportArgs te = ta[last ^ 1];
        te.inter = c->inter.v;
        te.next = c->pairq.v;
        te.next_cap = c->pairq.cap;
        HIPCHK(c, scopy(c, c->d_args, &te, sizeof(te), hipMemcpyHostToDevice));
        rc = lma_event_pass(c, me, true, n_ev, max_lanes, &ms[1], &launches[1]);
        if (rc) return rc;
        now[C_LNEXT] = read_ctr(c, C_LNEXT);
        c->n_out = read_ctr(c, C_OUT);
        c->n_inter = read_ctr(c, C_INTER);
    }
    *pcur = pb[last];
    *n_left = size_t(now[C_LNEXT]);
    return PTG_OK;
}

// The LMA loop. Lepton phase: the step kernel over the live electrons and
// positrons, then the emission-event kernel over the ones that emitted, until
// none is left in the field. Photon phase (photons tracked): the photon step
// kernel, then the pair-event kernel, until no photon is left; the pairs start
// the next lepton phase. Without photon tracking, photons are final state and
// those of the primary beam only pass the final cuts.
static int run_lma(ptg_ctx* c, TransportArgs a, double* ms, uint64_t* launches, uint64_t* gens, bool use_surr, uint64_t* surr_fb) {
    const size_t n = c->n_prim;
    const bool track = a.opt.tracking_photons != 0;
    // event kernels: up to 4096 waves (16 per CU)
    const size_t max_lanes = 262144;
    LmaArgs m{};
    m.tables = c->d_lma;
    m.sc = c->lsc;
    m.lp_ceil = c->d_lp_ceil;
    m.bw = c->d_bw;
    m.bsc = c->bsc;
    m.cp = c->d_cp;
    m.csc = c->csc;
    m.lcfa = c->d_lcfa;
    m.mode = !c->laser.circular ? (a.classical_mode ? LMA_LP_CLASSICAL : LMA_LP) : (a.classical_mode ? LMA_CP_CLASSICAL : LMA_CP);
    m.mixed = a.precision == PTG_PREC_MIXED ? 1 : 0;
    m.tables_f = c->d_lma_f;
    m.lasers = c->n_coll ? c->d_lasers : nullptr;
    m.coll = c->n_coll ? c->d_coll : nullptr;
    m.bandwidth = c->lma_bandwidth;
    m.pol_angle = c->pol_angle;
    m.sin_pa = std::sin(c->pol_angle);
    m.cos_pa = std::cos(c->pol_angle);
    m.ev_ctr = c->d_ctr + C_EV;
    m.ph_ctr = c->d_ctr + C_PH;
    m.pairs_ctr = c->d_ctr + C_PAIRS;
    m.photons_to_out = track ? 0 : 1;

    const unsigned long long zero = 0;
    auto zero_ctr = [&](int which) { return hipMemcpyAsync(c->d_ctr + which, &zero, sizeof(zero), hipMemcpyHostToDevice, c->stream); };
    HIPCHK(c, zero_ctr(C_NEXT));
    HIPCHK(c, zero_ctr(C_PH));
    DevBuf* cur = &c->prim;
    size_t n_cur = n;
    int li = 0;
    c->n_def = 0;
    for (;;) {
        // ---------------- lepton phase
        for (;;) {
        while (n_cur > 0) {
            if (c->launch_mode == PTG_LAUNCH_GRAPH && (cur == &c->lq[0] || cur == &c->lq[1])) {
                // after the phase's first generation: captured sequences
                int rc = lepton_sequence(c, a, m, track, use_surr, cur == &c->lq[0] ? 0 : 1, n_cur, max_lanes, ms, launches,
                                         gens, surr_fb, &cur, &n_cur, &li);
                if (rc) return rc;
                continue;
            }
            // each live lepton yields at most one final state or one event; each event
            // at most one photon and one continuing lepton
            HIPCHK(c, grow_buf(c->out, c->n_out + 2 * n_cur + 1024, c->n_out, c->stream));
            if (c->evq.cap < n_cur) HIPCHK(c, alloc_buf(c->evq, n_cur + n_cur / 4));
            int rc = ensure_ev_extra(c, c->evq.cap);
            if (rc) return rc;
            DevBuf& nxt = c->lq[li];
            if (&nxt == cur) return fail(c, PTG_E_STATE, "internal queue aliasing");
            if (nxt.cap < n_cur) HIPCHK(c, alloc_buf(nxt, n_cur + n_cur / 4));
            unsigned long long n_ph = read_ctr(c, track ? C_PH : C_NEXT);
            HIPCHK(c, grow_buf(c->phq, size_t(n_ph) + 2 * n_cur + 1024, size_t(n_ph), c->stream));
            HIPCHK(c, zero_ctr(C_WORK));
            HIPCHK(c, zero_ctr(C_EV));
            HIPCHK(c, zero_ctr(C_LNEXT));
            a.in = cur->v;
            a.n_in = uint32_t(n_cur);
            a.out = c->out.v;
            a.inter = c->inter.v;
            a.next = c->phq.v;  // photons among the inputs are handed on here
            a.next_ctr = c->d_ctr + (track ? C_PH : C_NEXT);
            a.next_cap = c->phq.cap;
            m.ev = c->evq.v;
            m.ev_cap = c->evq.cap;
            m.ev_k0 = c->d_ev_extra;
            m.ev_a = c->d_ev_extra + c->ev_extra_cap;
            m.ph = c->phq.v;
            m.ph_cap = c->phq.cap;
            HIPCHK(c, hipMemcpyAsync(c->d_args, &a, sizeof(a), hipMemcpyHostToDevice, c->stream));
            const size_t sb = size_t(c->step_block);
            int grid = int(std::min<size_t>(size_t(c->n_cu) * size_t(c->step_grid_per_cu), (n_cur + sb - 1) / sb));
            HIPCHK(c, hipEventRecord(c->ev0, c->stream));
            launch_step_lma(c->d_args, m, grid, c->stream, c->step_block);
            HIPCHK(c, hipGetLastError());
            HIPCHK(c, hipEventRecord(c->ev1, c->stream));
            unsigned long long n_ev = 0;
            HIPCHK(c, hipMemcpyAsync(&n_ev, c->d_ctr + C_EV, sizeof(n_ev), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            float t = 0.0f;
            (void)hipEventElapsedTime(&t, c->ev0, c->ev1);
            ms[0] += t;
            ++launches[0];
            if (n_ev > 0) {
                m.next = nxt.v;
                m.next_ctr = c->d_ctr + C_LNEXT;
                rc = use_surr ? lma_surrogate_pass(c, m, size_t(n_ev), max_lanes, &ms[1], &launches[1], surr_fb)
                              : lma_event_pass(c, m, false, size_t(n_ev), max_lanes, &ms[1], &launches[1]);
                if (rc) return rc;
            }
            unsigned long long now[C_N];
            HIPCHK(c, scopy(c, now, c->d_ctr, sizeof(now), hipMemcpyDeviceToHost));
            c->n_out = now[C_OUT];
            cur = &nxt;
            n_cur = size_t(now[C_LNEXT]);
            li ^= 1;
            ++*gens;
        }
        if (c->n_def == 0) break;
        // the events the learned sampler left to the exact kernels, in one pass; their
        // electrons continue in another lepton round
        {
            const size_t nd = c->n_def;
            HIPCHK(c, grow_buf(c->out, c->n_out + nd + 1024, c->n_out, c->stream));
            DevBuf& nxt = c->lq[li];
            if (nxt.cap < nd) HIPCHK(c, alloc_buf(nxt, nd + nd / 4));
            unsigned long long n_ph = read_ctr(c, track ? C_PH : C_NEXT);
            HIPCHK(c, grow_buf(c->phq, size_t(n_ph) + nd + 1024, size_t(n_ph), c->stream));
            HIPCHK(c, zero_ctr(C_LNEXT));
            a.out = c->out.v;
            a.next = c->phq.v;
            a.next_ctr = c->d_ctr + (track ? C_PH : C_NEXT);
            a.next_cap = c->phq.cap;
            HIPCHK(c, hipMemcpyAsync(c->d_args, &a, sizeof(a), hipMemcpyHostToDevice, c->stream));
            m.ph = c->phq.v;
            m.ph_cap = c->phq.cap;
            m.next = nxt.v;
            m.next_ctr = c->d_ctr + C_LNEXT;
            m.ev_k0 = c->d_def_extra;
            m.ev_a = c->d_def_extra + c->def_extra_cap;
            int rc = lma_event_pass(c, m, false, nd, max_lanes, &ms[1], &launches[1], nullptr, &c->defq);
            if (rc) return rc;
            c->n_def = 0;
            unsigned long long now[C_N];
            HIPCHK(c, scopy(c, now, c->d_ctr, sizeof(now), hipMemcpyDeviceToHost));
            c->n_out = now[C_OUT];
            cur = &nxt;
            n_cur = size_t(now[C_LNEXT]);
            li ^= 1;
            ++*gens;
        }
        if (n_cur == 0) break;
        }
        if (!track) break;
        // ---------------- photon phase
        size_t n_ph = size_t(read_ctr(c, C_PH));
        if (n_ph == 0) break;
        HIPCHK(c, zero_ctr(C_PAIRS));
        DevBuf* pcur = &c->phq;
        while (n_ph > 0) {
            if (c->launch_mode == PTG_LAUNCH_GRAPH) {
                int rc = photon_sequence(c, a, m, pcur == &c->phq ? 0 : 1, n_ph, max_lanes, ms, launches, gens, &pcur, &n_ph);
                if (rc) return rc;
                continue;
            }
            HIPCHK(c, grow_buf(c->out, c->n_out + n_ph + 1024, c->n_out, c->stream));
            if (a.opt.keep_decayed_photons) HIPCHK(c, grow_buf(c->inter, c->n_inter + n_ph + 1024, c->n_inter, c->stream));
            if (c->evq.cap < n_ph) HIPCHK(c, alloc_buf(c->evq, n_ph + n_ph / 4));
            int rc = ensure_ev_extra(c, c->evq.cap);
            if (rc) return rc;
            DevBuf& pnxt = pcur == &c->phq ? c->phq2 : c->phq;
            if (pnxt.cap < n_ph) HIPCHK(c, alloc_buf(pnxt, n_ph + n_ph / 4));
            unsigned long long n_pairs = read_ctr(c, C_PAIRS);
            HIPCHK(c, grow_buf(c->pairq, size_t(n_pairs) + 2 * n_ph + 1024, size_t(n_pairs), c->stream));
            HIPCHK(c, zero_ctr(C_WORK));
            HIPCHK(c, zero_ctr(C_EV));
            HIPCHK(c, zero_ctr(C_LNEXT));
            a.in = pcur->v;
            a.n_in = uint32_t(n_ph);
            a.out = c->out.v;
            a.inter = c->inter.v;
            a.next = c->pairq.v;  // (no leptons are queued with photons)
            a.next_ctr = c->d_ctr + C_PAIRS;
            a.next_cap = c->pairq.cap;
            m.ev = c->evq.v;
            m.ev_cap = c->evq.cap;
            m.ev_k0 = c->d_ev_extra;
            m.ev_a = c->d_ev_extra + c->ev_extra_cap;
            HIPCHK(c, hipMemcpyAsync(c->d_args, &a, sizeof(a), hipMemcpyHostToDevice, c->stream));
            int grid = int(std::min<size_t>(size_t(c->n_cu) * 4, (n_ph + 255) / 256));
            HIPCHK(c, hipEventRecord(c->ev0, c->stream));
            launch_photon_step_lma(c->d_args, m, grid, c->stream);
            HIPCHK(c, hipGetLastError());
            HIPCHK(c, hipEventRecord(c->ev1, c->stream));
            unsigned long long n_ev = 0;
            HIPCHK(c, hipMemcpyAsync(&n_ev, c->d_ctr + C_EV, sizeof(n_ev), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            float t = 0.0f;
            (void)hipEventElapsedTime(&t, c->ev0, c->ev1);
            ms[1] += t;
            ++launches[1];
            if (n_ev > 0) {
                m.next = pnxt.v;
                m.next_ctr = c->d_ctr + C_LNEXT;
                m.pairs = c->pairq.v;
                m.pairs_cap = c->pairq.cap;
                rc = lma_event_pass(c, m, true, size_t(n_ev), max_lanes, &ms[1], &launches[1]);
                if (rc) return rc;
            }
            unsigned long long now[C_N];
            HIPCHK(c, scopy(c, now, c->d_ctr, sizeof(now), hipMemcpyDeviceToHost));
            c->n_out = now[C_OUT];
            c->n_inter = now[C_INTER];
            pcur = &pnxt;
            n_ph = size_t(now[C_LNEXT]);
            ++*gens;
        }
        // the pairs start the next lepton phase
        n_cur = size_t(read_ctr(c, C_PAIRS));
        if (n_cur == 0) break;
        cur = &c->pairq;
        HIPCHK(c, zero_ctr(C_PH));
    }
    if (!track) {
        // photons of the primary beam: not tracked, final cuts only
        unsigned long long n_ph = read_ctr(c, C_NEXT);
        if (n_ph > 0) {
            TransportArgs b = a;
            b.laser.kind = K_FAST_PLANE_WAVE;  // the pass-through never evaluates fields
            b.opt.tracking_photons = 0;
            double pms = 0.0;
            uint64_t pl = 0;
            int rc = transport(c, b, true, c->phq, size_t(n_ph), c->q[0], &pms, &pl);
            if (rc) return rc;
        }
    }
    return PTG_OK;
}

int ptg_run(ptg_ctx* c, const ptg_collide_options* o, uint64_t seed, ptg_counters* cnt) {
    if (!c || !o) return fail(c, PTG_E_ARG, "null argument");
    if (!c->have_laser) return fail(c, PTG_E_STATE, "ptg_set_laser must be called first");
    if (c->n_prim == 0) return fail(c, PTG_E_STATE, "no primaries resident (ptg_load_primaries / ptg_build_beam)");
    int rc = validate_options(c, o);
    if (rc) return rc;
    if (o->surrogate) {
        if (!(c->laser.kind == K_PLANE_WAVE || c->laser.kind == K_FOCUSED_LASER))
            return fail(c, PTG_E_UNSUPPORTED, "the surrogate samples LMA emission (PlaneWave, FocusedLaser)");
        if (o->classical) return fail(c, PTG_E_UNSUPPORTED, "the surrogate samples quantum LMA emission (classical: false)");
        const int p = c->laser.circular ? 1 : 0;
        if (!c->surr_loaded[p]) return fail(c, PTG_E_STATE, "no surrogate loaded for this polarization (ptg_surrogate_load)");
        if (!c->surr_valid[p]) return fail(c, PTG_E_STATE, "the loaded surrogate did not pass its validation");
    }
    if (o->precision == PTG_PREC_MIXED) {
        if (!(c->laser.kind == K_PLANE_WAVE || c->laser.kind == K_FOCUSED_LASER))
            return fail(c, PTG_E_UNSUPPORTED, "mixed precision is implemented for the LMA fields (PlaneWave, FocusedLaser)");
        if (o->classical)
            return fail(c, PTG_E_UNSUPPORTED, "mixed precision is implemented for quantum LMA emission (classical: false)");
    }
    if (c->laser.kind == K_PLANE_WAVE || c->laser.kind == K_FOCUSED_LASER) {
        // classical + rr + gaunt_factor is the modified Landau-Lifshitz equation,
        // which the LMA fields refuse (plane_wave.rs:214,228,236; focused_laser.rs:223,237,245)
        if (o->classical && o->rr && o->gaunt_factor)
            return fail(c, PTG_E_ARG, "Gaunt factor correction is unavailable in LMA mode");
        if (o->tracking_photons && !c->laser.circular && !c->d_bw)
            return fail(c, PTG_E_UNSUPPORTED, "LMA pair creation needs the regenerated Breit-Wheeler mid-range table "
                                              "(tables/gen_bw_tables.py, then tables/extract.py)");
        if (c->laser.circular && !c->d_cp)
            return fail(c, PTG_E_UNSUPPORTED, "circularly polarised LMA emission needs the regenerated CP CDF table "
                                              "(tables/gen_cp_tables.py, then tables/extract.py)");

    }
    HIPCHK(c, hipSetDevice(c->device));
    const size_t n = c->n_prim;

    TransportArgs a{};
    a.laser = c->laser;
    a.sc = c->sc;
    a.dt = c->laser.dt * o->dt_multiplier;
    a.opt.rate_increase = o->rate_increase;
    a.opt.t_stop = o->t_stop;
    a.opt.min_energy = o->min_energy;
    a.opt.cos_max_angle = std::cos(o->max_angle);
    a.opt.tracking_photons = o->tracking_photons;
    a.opt.tracking_secondaries = o->tracking_secondaries;
    a.opt.keep_decayed_photons = o->keep_decayed_photons;
    a.opt.pol_resolved = o->pol_resolved;
    a.opt.rotate_stokes_pars = o->rotate_stokes_pars;
    a.opt.discard_bg_e = o->discard_bg_e;
    a.opt.discard_bg_ph = o->discard_bg_ph;
    a.opt.multiplicity_on = o->multiplicity >= 0;
    a.rr_mode = (o->classical && o->rr) ? (o->gaunt_factor ? 2 : 1) : 0;
    a.classical_mode = o->classical && !o->gaunt_factor;
    a.recoil = !o->classical && o->rr;
    a.precision = o->precision;
    a.key = key_from_seed(seed);
    a.step_cap = 1u << 22;
    {
        // peak |E| + c|B| of the pulse, with a factor 1000 of margin for the
        // focusing corrections: below ff_kt_max a photon's chi stays under 0.01
        double e_peak = c->laser.kind == K_FAST_FOCUSED_LASER ? 2.0 * c->laser.E0 : 2.0 * c->laser.amplitude;
        a.ff_kt_max = e_peak > 0.0 ? 0.01 * E_CRIT / (1000.0 * e_peak) : 0.0;
        a.ff_after = 4096;
    }
    a.lcfa_tables = c->d_lcfa;
    a.work_ctr = reinterpret_cast<uint32_t*>(c->d_ctr + C_WORK);
    a.out_ctr = c->d_ctr + C_OUT;
    a.inter_ctr = c->d_ctr + C_INTER;
    a.next_ctr = c->d_ctr + C_NEXT;
    a.id_ctr = c->d_ctr + C_ID;
    a.id_base = n;
    a.prim_n0 = c->d_n0;
    a.n_prim = uint32_t(n);
    a.stats = c->d_ctr + C_STATS;
    a.flags = reinterpret_cast<uint32_t*>(c->d_ctr + C_FLAGS);

    HIPCHK(c, hipMemsetAsync(c->d_ctr, 0, C_N * sizeof(unsigned long long), c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->n_out = 0;
    c->n_inter = 0;
    c->graph_launches = 0;
    c->graph_stalls = 0;

    double ms[2] = {0.0, 0.0};  // lepton, photon kernels
    uint64_t launches[2] = {0, 0};
    uint64_t gens = 0;
    uint64_t surr_fb = 0;
    if (c->laser.kind == K_PLANE_WAVE || c->laser.kind == K_FOCUSED_LASER) {
        rc = run_lma(c, a, ms, launches, &gens, o->surrogate != 0, &surr_fb);
        if (rc) return rc;
    } else {
        // Generation loop: primaries and pairs are electrons/positrons (or photons,
        // for photon beams); every kernel consumes one queue and fills the other.
        DevBuf* cur = &c->prim;
        size_t n_cur = n;
        bool photon_stage = false;  // leptons first; each stage forwards the other kind
        int qi = 0;
        while (n_cur > 0) {
            DevBuf& next = c->q[qi];
            if (&next == cur) return fail(c, PTG_E_STATE, "internal queue aliasing");
            rc = transport(c, a, photon_stage, *cur, n_cur, next, &ms[photon_stage], &launches[photon_stage]);
            if (rc) return rc;
            unsigned long long nn = 0;
            HIPCHK(c, scopy(c, &nn, c->d_ctr + C_NEXT, sizeof(nn), hipMemcpyDeviceToHost));
            cur = &next;
            n_cur = size_t(nn);
            qi ^= 1;
            photon_stage = !photon_stage;
            ++gens;
        }
    }

    if (a.opt.multiplicity_on && c->n_out) {
        // count each shower's final state now that it is complete (a repeated
        // transport launch cannot count a particle twice), then keep the showers
        // with multiplicity + 1 particles
        if (c->shower_cap < n) {
            if (c->d_shower_count) (void)hipFree(c->d_shower_count);
            c->d_shower_count = nullptr;
            c->shower_cap = 0;
            HIPCHK(c, hipMalloc(&c->d_shower_count, n * sizeof(uint32_t)));
            c->shower_cap = n;
        }
        HIPCHK(c, hipMemsetAsync(c->d_shower_count, 0, n * sizeof(uint32_t), c->stream));
        launch_count_showers(c->out.v, c->n_out, c->d_shower_count, c->stream);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, grow_buf(c->tmp, c->n_out, 0, c->stream));
        unsigned long long zero = 0;
        HIPCHK(c, scopy(c, c->d_ctr + C_WORK, &zero, sizeof(zero), hipMemcpyHostToDevice));
        launch_select_multiplicity(c->out.v, c->n_out, c->tmp.v, c->d_ctr + C_WORK, c->d_shower_count,
                                   uint32_t(o->multiplicity + 1), c->stream);
        HIPCHK(c, hipGetLastError());
        unsigned long long kept = 0;
        HIPCHK(c, scopy(c, &kept, c->d_ctr + C_WORK, sizeof(kept), hipMemcpyDeviceToHost));
        std::swap(c->out, c->tmp);
        c->n_out = kept;
    }

    unsigned long long fin[C_N];
    HIPCHK(c, scopy(c, fin, c->d_ctr, sizeof(fin), hipMemcpyDeviceToHost));
    if (cnt) {
        cnt->lepton_steps = fin[C_STATS + STAT_LEPTON_STEPS];
        cnt->photon_steps = fin[C_STATS + STAT_PHOTON_STEPS];
        cnt->emissions = fin[C_STATS + STAT_EMISSIONS];
        cnt->pair_events = fin[C_STATS + STAT_PAIRS];
        cnt->rejection_trials = fin[C_STATS + STAT_TRIALS];
        cnt->n_secondaries = c->n_out;
        cnt->n_intermediates = c->n_inter;
        cnt->generations = gens;
        cnt->error_flags = uint32_t(fin[C_FLAGS]);
        cnt->kernel_ms = ms[0] + ms[1];
        cnt->lepton_ms = ms[0];
        cnt->photon_ms = ms[1];
        cnt->lepton_launches = launches[0];
        cnt->photon_launches = launches[1];
        cnt->photon_steps_closed_form = fin[C_STATS + STAT_PHOTON_STEPS_CF];
        cnt->surrogate_events = fin[C_STATS + STAT_SURROGATE];
        cnt->surrogate_fallbacks = surr_fb + fin[C_STATS + STAT_SURR_FB];
        cnt->graph_launches = c->graph_launches;
        cnt->graph_stalls = c->graph_stalls;
        cnt->ceiling_restarts = fin[C_STATS + STAT_CEIL_RESTART];
    }
    return PTG_OK;
}

int ptg_collide(ptg_ctx* c, const ptg_particles* p, size_t n, const ptg_collide_options* o, uint64_t seed, ptg_counters* cnt) {
    int rc = ptg_load_primaries(c, p, n);
    if (rc) return rc;
    return ptg_run(c, o, seed, cnt);
}

static int fetch_buf(ptg_ctx* c, const DevBuf& b, size_t n, ptg_particles* p) {
    if (!p || n == 0) return PTG_OK;
    std::vector<double> col(n);
    auto get4 = [&](int first, double* dst) -> int {
        if (!dst) return PTG_OK;
        for (int k = 0; k < 4; ++k) {
            HIPCHK(c, scopy(c, col.data(), b.v.d[first + k], n * 8, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < n; ++i) dst[4 * i + k] = col[i];
        }
        return PTG_OK;
    };
    auto get1 = [&](int k, double* dst) -> int {
        if (dst) HIPCHK(c, scopy(c, dst, b.v.d[k], n * 8, hipMemcpyDeviceToHost));
        return PTG_OK;
    };
    int rc = 0;
    if ((rc = get4(P_RT, p->r)) || (rc = get4(P_UT, p->u)) || (rc = get4(P_R0T, p->r0)) || (rc = get4(P_U0T, p->u0)) ||
        (rc = get4(P_SI, p->pol)) || (rc = get1(P_PAYLOAD, p->payload)) || (rc = get1(P_PCHI, p->parent_chi)) ||
        (rc = get1(P_ICOUNT, p->interaction_count)) || (rc = get1(P_WORK, p->work)) || (rc = get1(P_WEIGHT, p->weight)))
        return rc;
    if (p->id) HIPCHK(c, scopy(c, p->id, b.v.id, n * 8, hipMemcpyDeviceToHost));
    if (p->parent_id) HIPCHK(c, scopy(c, p->parent_id, b.v.parent_id, n * 8, hipMemcpyDeviceToHost));
    if (p->species) {
        std::vector<uint32_t> m(n);
        HIPCHK(c, scopy(c, m.data(), b.v.meta, n * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) p->species[i] = uint8_t(m[i] & 3u);
    }
    return PTG_OK;
}

int ptg_fetch(ptg_ctx* c, ptg_particles* sec, ptg_particles* inter) {
    if (!c) return PTG_E_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    int rc = fetch_buf(c, c->out, c->n_out, sec);
    if (rc) return rc;
    return fetch_buf(c, c->inter, c->n_inter, inter);
}

int ptg_fetch_showers(ptg_ctx* c, uint32_t* shower) {
    if (!c || !shower) return fail(c, PTG_E_ARG, "null argument");
    if (c->n_out == 0) return PTG_OK;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, scopy(c, shower, c->out.v.shower, c->n_out * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return PTG_OK;
}

int ptg_push(ptg_ctx* c, const uint8_t* species, double* r, double* u, size_t n, uint32_t n_steps, const ptg_collide_options* o,
             uint32_t* steps_taken, double* work) {
    if (!c || !species || !r || !u || !o) return fail(c, PTG_E_ARG, "null argument");
    if (!c->have_laser) return fail(c, PTG_E_STATE, "ptg_set_laser must be called first");
    if (n == 0) return PTG_OK;
    HIPCHK(c, hipSetDevice(c->device));
    uint8_t* ds = nullptr;
    double *dr = nullptr, *du = nullptr, *dw = nullptr;
    uint32_t* dk = nullptr;
    HIPCHK(c, hipMalloc(&ds, n));
    HIPCHK(c, hipMalloc(&dr, n * 32));
    HIPCHK(c, hipMalloc(&du, n * 32));
    HIPCHK(c, hipMalloc(&dk, n * 4));
    HIPCHK(c, hipMalloc(&dw, n * 8));
    HIPCHK(c, scopy(c, ds, species, n, hipMemcpyHostToDevice));
    HIPCHK(c, scopy(c, dr, r, n * 32, hipMemcpyHostToDevice));
    HIPCHK(c, scopy(c, du, u, n * 32, hipMemcpyHostToDevice));
    HIPCHK(c, sfill(c, dw, 0, n * 8));
    int rr_mode = (o->classical && o->rr) ? (o->gaunt_factor ? 2 : 1) : 0;
    double dt = c->laser.dt * o->dt_multiplier;
    if (c->laser.kind == K_PLANE_WAVE || c->laser.kind == K_FOCUSED_LASER) {
        if (rr_mode == 2) return fail(c, PTG_E_ARG, "Gaunt factor correction is unavailable in LMA mode");
        launch_push_lma(c->laser, dt, rr_mode, ds, dr, du, uint32_t(n), n_steps, dk, dw, c->stream);
    } else {
        launch_push_lcfa(c->laser, dt, rr_mode, ds, dr, du, uint32_t(n), n_steps, dk, dw, c->stream);
    }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, scopy(c, r, dr, n * 32, hipMemcpyDeviceToHost));
    HIPCHK(c, scopy(c, u, du, n * 32, hipMemcpyDeviceToHost));
    if (steps_taken) HIPCHK(c, scopy(c, steps_taken, dk, n * 4, hipMemcpyDeviceToHost));
    if (work) HIPCHK(c, scopy(c, work, dw, n * 8, hipMemcpyDeviceToHost));
    (void)hipFree(ds); (void)hipFree(dr); (void)hipFree(du); (void)hipFree(dk); (void)hipFree(dw);
    return PTG_OK;
}

// The range cut of `cut` and the unit scale of `s` (scale 0 means 1).
static int hist_cut_scale(ptg_ctx* c, HistArgs& h, const ptg_hist_spec* cut, const ptg_hist_spec* s) {
    if (cut->weight_fn < PTG_W_NUMBER || cut->weight_fn > PTG_W_POL_Y) return fail(c, PTG_E_ARG, "unknown weight function");
    if (cut->cut && (cut->cut_quantity < 0 || cut->cut_quantity >= PTG_Q_COUNT))
        return fail(c, PTG_E_ARG, "unknown cut quantity");
    if (!(std::isfinite(s->scale) && s->scale >= 0.0)) return fail(c, PTG_E_ARG, "scale must be finite and >= 0");
    h.cut = cut->cut ? 1 : 0;
    h.cq = cut->cut ? cut->cut_quantity : 0;
    h.cmin = cut->cut_min;
    h.cmax = cut->cut_max;
    h.scale = s->scale == 0.0 ? 1.0 : s->scale;
    return PTG_OK;
}

// Histogram::generate_1d over the resident final state into the caller's host
// buffers, with the bins accumulator at d_bins (nbins + 1 doubles) and the two
// range keys at d_keys on the device. Allocates nothing; clears with kernels.
static int histogram_into(ptg_ctx* c, ptg_hist_spec* s, double* bins_out, double* total_weight, double* d_bins,
                          unsigned long long* d_keys) {
    HistArgs h{};
    // species 3: the intermediate (decayed) photons kept by keep_decayed_photons
    h.p = s->species == 3 ? c->inter.v : c->out.v;
    h.n = s->species == 3 ? c->n_inter : c->n_out;
    h.species = s->species == 3 ? 2 : s->species;
    h.q = s->quantity;
    h.wfn = s->weight_fn;
    h.log_scale = s->log_scale;
    h.nbins = s->nbins;
    if (int rc = hist_cut_scale(c, h, s, s)) return rc;
    int grid = std::max(1, std::min(c->n_cu * 4, int((h.n + 255) / 256)));
    double gmin = s->min, gmax = s->max;
    if (!s->fixed_range) {
        unsigned long long mm[2] = {~0ull, 0ull};
        launch_fill_u64(d_keys, 1, ~0ull, c->stream);
        launch_fill_u64(d_keys + 1, 1, 0ull, c->stream);
        if (h.n) launch_hist_minmax(h, d_keys, grid, c->stream);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, scopy(c, mm, d_keys, sizeof(mm), hipMemcpyDeviceToHost));
        if (mm[0] == ~0ull) {  // nothing to bin (reference returns None)
            for (uint32_t b = 0; b < s->nbins; ++b) bins_out[b] = 0.0;
            if (total_weight) *total_weight = 0.0;
            return fail(c, PTG_E_DOMAIN, "no particles of this species to histogram");
        }
        gmin = host_order_key_inv(mm[0]);
        gmax = host_order_key_inv(mm[1]);
    }
    uint32_t nbins = gmin == gmax ? 1u : s->nbins;  // hgram.rs:175-177
    if (nbins != s->nbins) for (uint32_t b = 0; b < s->nbins; ++b) bins_out[b] = 0.0;
    h.nbins = nbins;
    h.gmin = gmin;
    h.bin_vol = gmin == gmax ? 1.0 : (gmax - gmin) / double(nbins);
    launch_fill_u64(reinterpret_cast<unsigned long long*>(d_bins), nbins + 1, 0ull, c->stream);  // +0.0
    if (h.n) launch_hist_bin(h, d_bins, d_bins + nbins, grid, c->stream);
    HIPCHK(c, hipGetLastError());
    double tot = 0.0;
    HIPCHK(c, hipMemcpyAsync(bins_out, d_bins, nbins * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(&tot, d_bins + nbins, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (total_weight) *total_weight = tot;
    s->min = gmin;
    s->max = gmax;
    return PTG_OK;
}

static int check_hist_spec(ptg_ctx* c, const ptg_hist_spec* s, uint32_t max_bins) {
    if (s->nbins == 0 || s->nbins > max_bins)
        return fail(c, PTG_E_ARG, "nbins must be in [1, " + std::to_string(max_bins) + "]");
    if (s->quantity < 0 || s->quantity >= PTG_Q_COUNT) return fail(c, PTG_E_ARG, "unknown quantity");
    if (s->species < 0 || s->species > 3) return fail(c, PTG_E_ARG, "bad species");
    return PTG_OK;
}

int ptg_histogram(ptg_ctx* c, ptg_hist_spec* s, double* bins_out, double* total_weight) {
    if (!c || !s || !bins_out) return fail(c, PTG_E_ARG, "null argument");
    if (int rc = check_hist_spec(c, s, 1u << 20)) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    if (c->hist_cap < size_t(s->nbins) + 3) {
        if (c->d_hist) (void)hipFree(c->d_hist);
        c->d_hist = nullptr;
        c->hist_cap = 0;
        const size_t cap = std::max<size_t>(size_t(s->nbins) + 3, 8192 + 3);
        HIPCHK(c, hipMalloc(&c->d_hist, cap * sizeof(double)));
        c->hist_cap = cap;
    }
    return histogram_into(c, s, bins_out, total_weight, c->d_hist,
                          reinterpret_cast<unsigned long long*>(c->d_hist + c->hist_cap - 2));
}

int ptg_histogram_noalloc(ptg_ctx* c, ptg_hist_spec* s, double* bins_out, double* total_weight) {
    if (!c || !s || !bins_out) return fail(c, PTG_E_ARG, "null argument");
    if (s->nbins > PTG_NOALLOC_MAX_BINS)
        return fail(c, PTG_E_ARG, "ptg_histogram_noalloc bins at most PTG_NOALLOC_MAX_BINS (" +
                                      std::to_string(PTG_NOALLOC_MAX_BINS) + ") values; use ptg_histogram for more");
    if (int rc = check_hist_spec(c, s, PTG_NOALLOC_MAX_BINS)) return rc;
    if (!c->d_noalloc) return fail(c, PTG_E_STATE, "no scratch address for this device");
    HIPCHK(c, hipSetDevice(c->device));
    return histogram_into(c, s, bins_out, total_weight, reinterpret_cast<double*>(c->d_noalloc),
                          c->d_noalloc + PTG_NOALLOC_MAX_BINS + 1);
}

int ptg_finish(ptg_ctx* c, int stop_mode, double t_stop, int beam_basis, double theta, double phi, double* t_common) {
    if (!c) return fail(c, PTG_E_ARG, "null argument");
    if (stop_mode < 0 || stop_mode > 2) return fail(c, PTG_E_ARG, "stop_mode must be 0, 1 or 2");
    if (stop_mode == 2 && !std::isfinite(t_stop)) return fail(c, PTG_E_ARG, "t_stop must be finite");
    HIPCHK(c, hipSetDevice(c->device));
    double ct = INFINITY;
    if (stop_mode == 1 && c->n_out) {
        unsigned long long* d = nullptr;
        HIPCHK(c, hipMalloc(&d, sizeof(unsigned long long)));
        unsigned long long k = 0ull;
        hipError_t e = scopy(c, d, &k, sizeof(k), hipMemcpyHostToDevice);
        if (e == hipSuccess) { launch_max_ct(c->out.v, c->n_out, d, c->stream); e = hipGetLastError(); }
        if (e == hipSuccess) e = scopy(c, &k, d, sizeof(k), hipMemcpyDeviceToHost);
        (void)hipFree(d);
        HIPCHK(c, e);
        ct = host_order_key_inv(k);
    } else if (stop_mode == 1) {
        ct = -INFINITY;  // no particles: the reference's NEG_INFINITY start
    } else if (stop_mode == 2) {
        ct = t_stop * C_LIGHT;
    }
    // beam centroid (-cos phi sin theta, -sin theta sin phi, -cos theta) onto +z (particle/mod.rs:256-272)
    double vangle = 0.0, hangle = PI_D - theta;
    if (phi != 0.0 && theta != 0.0) {
        const double cos_v = std::hypot(std::cos(theta), std::sin(theta) * std::cos(phi));
        vangle = std::copysign(std::acos(cos_v), std::sin(theta) * std::sin(phi));
        hangle = std::copysign(std::acos(-std::cos(theta) / cos_v), std::cos(phi) * std::sin(theta));
    }
    launch_finish(c->out.v, c->n_out, stop_mode != 0 && c->n_out, ct, beam_basis ? 1 : 0, hangle, vangle, c->stream);
    HIPCHK(c, hipGetLastError());
    // the intermediate photons take the basis change but not the drift: the
    // reference drifts bunch.iter_mut() and rotates bunch.iter_all_mut()
    // (src/main.rs:1150-1196, particle/shower.rs:64-69)
    if (beam_basis && c->n_inter) {
        launch_finish(c->inter.v, c->n_inter, 0, ct, 1, hangle, vangle, c->stream);
        HIPCHK(c, hipGetLastError());
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (t_common) *t_common = stop_mode == 0 ? INFINITY : ct / C_LIGHT;
    return PTG_OK;
}

int ptg_stats(ptg_ctx* c, const ptg_stats_spec* s, double* out) {
    if (!c || !s || !out) return fail(c, PTG_E_ARG, "null argument");
    for (int q : {s->variable, s->weight, s->cut ? s->cut_quantity : 0})
        if (q < 0 || q >= PTG_Q_COUNT) return fail(c, PTG_E_ARG, "unknown quantity");
    if (s->species < 0 || s->species > 3) return fail(c, PTG_E_ARG, "bad species");
    HIPCHK(c, hipSetDevice(c->device));
    StatArgs a{};
    a.p = s->species == 3 ? c->inter.v : c->out.v;  // 3: the intermediate photons
    a.n = s->species == 3 ? c->n_inter : c->n_out;
    a.species = s->species == 3 ? 2 : s->species;
    a.qv = s->variable;
    a.qw = s->weight;
    a.cut = s->cut ? 1 : 0;
    a.qf = s->cut ? s->cut_quantity : 0;
    a.fmin = s->cut_min;
    a.fmax = s->cut_max;
    a.shift = s->shift;
    // 7 sums (slots 0-6) then the two order keys (slots 8, 9)
    unsigned long long* d = nullptr;
    HIPCHK(c, hipMalloc(&d, 10 * sizeof(unsigned long long)));
    unsigned long long init[10] = {0, 0, 0, 0, 0, 0, 0, 0, ~0ull, 0ull};
    unsigned long long h[10];
    hipError_t e = scopy(c, d, init, sizeof(init), hipMemcpyHostToDevice);
    int grid = std::max(1, std::min(c->n_cu * 4, int((a.n + 255) / 256)));
    if (e == hipSuccess && a.n) {
        launch_stats(a, reinterpret_cast<double*>(d), d + 8, grid, c->stream);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = scopy(c, h, d, sizeof(h), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    HIPCHK(c, e);
    double sums[7];
    memcpy(sums, h, sizeof(sums));
    out[0] = sums[0]; out[1] = sums[1]; out[2] = sums[2];
    out[3] = h[8] == ~0ull ? INFINITY : host_order_key_inv(h[8]);
    out[4] = h[8] == ~0ull ? -INFINITY : host_order_key_inv(h[9]);
    out[5] = sums[3]; out[6] = sums[4]; out[7] = sums[5]; out[8] = sums[6];
    return PTG_OK;
}

int ptg_histogram_2d(ptg_ctx* c, ptg_hist_spec* sx, ptg_hist_spec* sy, double* bins_out, double* total_weight) {
    if (!c || !sx || !sy || !bins_out) return fail(c, PTG_E_ARG, "null argument");
    for (const ptg_hist_spec* s : {sx, sy}) {
        if (s->nbins == 0 || s->nbins > 4096) return fail(c, PTG_E_ARG, "nbins must be in [1, 4096] per axis");
        if (s->quantity < 0 || s->quantity >= PTG_Q_COUNT) return fail(c, PTG_E_ARG, "unknown quantity");
    }
    HIPCHK(c, hipSetDevice(c->device));
    HistArgs h[2];
    const ptg_hist_spec* sp[2] = {sx, sy};
    const uint64_t n_src = sx->species == 3 ? c->n_inter : c->n_out;
    int grid = std::max(1, std::min(c->n_cu * 4, int((n_src + 255) / 256)));
    for (int a = 0; a < 2; ++a) {
        h[a] = HistArgs{};
        h[a].p = sx->species == 3 ? c->inter.v : c->out.v;  // 3: the intermediate photons
        h[a].n = sx->species == 3 ? c->n_inter : c->n_out;
        h[a].species = sx->species == 3 ? 2 : sx->species;  // one filter and one weight for both axes (hgram.rs:326-331)
        h[a].q = sp[a]->quantity;
        h[a].wfn = sx->weight_fn;
        h[a].log_scale = sp[a]->log_scale;
        if (int rc = hist_cut_scale(c, h[a], sx, sp[a])) return rc;
        double gmin = sp[a]->min, gmax = sp[a]->max;
        if (!sp[a]->fixed_range) {
            unsigned long long mm[2] = {~0ull, 0ull};
            HIPCHK(c, scopy(c, c->d_ctr + C_WORK, mm, sizeof(mm), hipMemcpyHostToDevice));
            if (h[a].n) launch_hist_minmax(h[a], c->d_ctr + C_WORK, grid, c->stream);
            HIPCHK(c, hipGetLastError());
            HIPCHK(c, scopy(c, mm, c->d_ctr + C_WORK, sizeof(mm), hipMemcpyDeviceToHost));
            if (mm[0] == ~0ull) {
                if (total_weight) *total_weight = 0.0;
                return fail(c, PTG_E_DOMAIN, "no particles of this species to histogram");
            }
            gmin = host_order_key_inv(mm[0]);
            gmax = host_order_key_inv(mm[1]);
        }
        h[a].nbins = gmin == gmax ? 1u : sp[a]->nbins;  // number_of_bins (hgram.rs:171-186)
        h[a].gmin = gmin;
        h[a].bin_vol = gmin == gmax ? 0.0 : (gmax - gmin) / double(h[a].nbins);  // bin size 0 for a degenerate axis
        const_cast<ptg_hist_spec*>(sp[a])->min = gmin;
        const_cast<ptg_hist_spec*>(sp[a])->max = gmax;
    }
    const size_t nb = size_t(h[0].nbins) * h[1].nbins;
    for (size_t b = 0; b < size_t(sx->nbins) * sy->nbins; ++b) bins_out[b] = 0.0;
    double* d_bins = nullptr;
    HIPCHK(c, hipMalloc(&d_bins, (nb + 1) * sizeof(double)));
    HIPCHK(c, hipMemsetAsync(d_bins, 0, (nb + 1) * sizeof(double), c->stream));
    if (c->n_out) launch_hist_bin2(h[0], h[1], d_bins, d_bins + nb, grid, c->stream);
    HIPCHK(c, hipGetLastError());
    std::vector<double> hb(nb + 1);
    HIPCHK(c, hipMemcpyAsync(hb.data(), d_bins, (nb + 1) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    (void)hipFree(d_bins);
    for (size_t b = 0; b < nb; ++b) bins_out[b] = hb[b];
    if (total_weight) *total_weight = hb[nb];
    return PTG_OK;
}

int ptg_eval(ptg_ctx* c, int fn, const double* in, size_t n, size_t in_w, double* out, size_t out_w) {
    if (!c || !in || !out) return fail(c, PTG_E_ARG, "null argument");
    if (fn < 0 || fn >= PTG_FN_COUNT) return fail(c, PTG_E_ARG, "unknown function id");
    static const int need_in[] = {2, 5, 1, 1, 4, 1, 3, 12, 2, 3, 4, 4, 2, 4, 6, 2, 3, 4, 4, 5, 2, 2, 2, 4, 2, 3, 4, 7, 3};
    static const int need_out[] = {1, 4, 2, 4, 7, 1, 1, 4, 2, 1, 5, 6, 3, 1, 4, 3, 2, 1, 2, 7, 2, 1, 1, 6, 3, 2, 2, 4, 2};
    static_assert(sizeof(need_in) / sizeof(int) == PTG_FN_COUNT && sizeof(need_out) / sizeof(int) == PTG_FN_COUNT, "probe widths");
    if (int(in_w) < need_in[fn] || int(out_w) < need_out[fn]) return fail(c, PTG_E_ARG, "row width too small for this function");
    if (fn == PTG_FN_FIELDS && !c->have_laser) return fail(c, PTG_E_STATE, "no laser set");
    if (n == 0) return PTG_OK;
    if (fn >= PTG_FN_BW_RATE && fn <= PTG_FN_BW_SAMPLE && !c->d_bw) return fail(c, PTG_E_UNSUPPORTED, "Breit-Wheeler tables not loaded");
    if (((fn >= PTG_FN_CP_RATE && fn <= PTG_FN_CP_SAMPLE) || (fn >= PTG_FN_CPBW_RATE && fn <= PTG_FN_CPBW_SAMPLE)) && !c->d_cp)
        return fail(c, PTG_E_UNSUPPORTED, "CP tables not loaded");
    if (fn >= PTG_FN_CPBW_RATE && fn <= PTG_FN_CPBW_SAMPLE) {
        for (size_t i = 0; i < n; ++i) {
            const double* x = in + i * in_w;
            double a = fn == PTG_FN_CPBW_PARTIAL_RATE ? x[1] : x[0], eta = fn == PTG_FN_CPBW_PARTIAL_RATE ? x[2] : x[1];
            if (!(a > 0.0 && a < 100.0 && eta > 0.0 && eta < 100.0))
                return fail(c, PTG_E_ARG, "CP Breit-Wheeler probes need 0 < a < 100 and 0 < eta < 100");
            if (fn == PTG_FN_CPBW_PARTIAL_RATE && !(x[0] >= 1.0 && x[0] <= 1.0e6 && double(int(x[0])) == x[0]))
                return fail(c, PTG_E_ARG, "CP Breit-Wheeler partial-rate probe needs an integer 1 <= n <= 1e6");
            if (fn == PTG_FN_CPBW_SAMPLE && !(x[6] == 0.0 || (x[6] >= 1.0 && x[6] <= 1.0e6)))
                return fail(c, PTG_E_ARG, "CP Breit-Wheeler sample probe needs fixed n = 0 or 1 <= n <= 1e6");
            if (fn != PTG_FN_CPBW_PARTIAL_RATE && a < 0.3 && 2.0 * (0.25 + a * a) / eta > 1.0e5)
                return fail(c, PTG_E_ARG, "CP Breit-Wheeler probe: more than 1e5 harmonics to sum");
        }
    }
    if (fn == PTG_FN_LP_CEILING) {
        for (size_t i = 0; i < n; ++i) {
            const double* x = in + i * in_w;
            if (!(x[0] >= 2.0 && x[0] <= 2.0e5 && double(int(x[0])) == x[0] && x[1] > 0.0 && x[1] <= 100.0 && x[2] > 0.0 &&
                  x[2] <= 100.0))
                return fail(c, PTG_E_ARG, "LP ceiling probe needs an integer 2 <= n <= 2e5, 0 < a <= 100, 0 < eta <= 100");
        }
    }
    if (fn == PTG_FN_BESSEL_J || fn == PTG_FN_CP_PARTIAL_RATE || fn == PTG_FN_CP_SAMPLE) {
        for (size_t i = 0; i < n; ++i) {
            const double* x = in + i * in_w;
            if (fn == PTG_FN_BESSEL_J && !(x[0] >= 1.0 && x[0] <= 1.0e7 && double(int(x[0])) == x[0] && x[1] >= 0.0))
                return fail(c, PTG_E_ARG, "Bessel probe needs an integer 1 <= n <= 1e7 and x >= 0");
            if (fn == PTG_FN_CP_PARTIAL_RATE && !(x[0] >= 1.0 && x[0] <= 1.0e6 && double(int(x[0])) == x[0] && x[1] > 0.0 && x[2] > 0.0))
                return fail(c, PTG_E_ARG, "CP partial-rate probe needs an integer 1 <= n <= 1e6, a > 0, eta > 0");
            if (fn == PTG_FN_CP_SAMPLE && !(x[3] >= 0.0 && x[3] <= 1.0e6 && x[0] > 0.0 && x[1] > 0.0))
                return fail(c, PTG_E_ARG, "CP sample probe needs a > 0, eta > 0 and 0 <= fixed n <= 1e6");
        }
    }
    if (fn >= PTG_FN_LP_RATE) {
        HIPCHK(c, hipSetDevice(c->device));
        if (fn == PTG_FN_BW_SAMPLE) {
            for (size_t i = 0; i < n; ++i)
                if (!(in[i * in_w + 5] >= 0.0 && in[i * in_w + 5] <= 2.0e5))
                    return fail(c, PTG_E_ARG, "BW sample probe needs 0 <= fixed n <= 2e5");
        }
        size_t len = 0;  // orders of the recurrence probe's per-row workspace
        if (fn == PTG_FN_DOUBLE_BESSEL || fn == PTG_FN_LP_SAMPLE || fn == PTG_FN_LP_CL_SAMPLE || fn == PTG_FN_LP_CL_CEILING) {
            for (size_t i = 0; i < n; ++i) {
                const double* x = in + i * in_w;
                if (fn == PTG_FN_DOUBLE_BESSEL) {
                    // the recurrence's workspace is sized for x <= n sqrt 2, y <= n / 2 (double_bessel.rs:39-76)
                    const int nn = int(x[0]);
                    if (!(x[0] >= 3.0 && x[0] <= 2.0e5) || double(nn) != x[0])
                        return fail(c, PTG_E_ARG, "J(n, x, y) probe needs an integer 3 <= n <= 2e5");
                    if (!(std::fabs(x[1]) <= nn * 1.4142135623730951 && x[2] >= 0.0 && x[2] <= 0.5 * nn))
                        return fail(c, PTG_E_ARG, "J(n, x, y) probe needs |x| <= n sqrt 2 and 0 <= y <= n / 2");
                    if (!(x[3] == 0.0 || x[3] == 1.0 || x[3] == 2.0 || x[3] == 3.0 || x[3] == 4.0))
                        return fail(c, PTG_E_ARG, "J(n, x, y) probe mode is 0, 1, 2, 3 or 4");
                    if (x[3] == 1.0 || x[3] == 3.0) len = std::max(len, size_t(dbessel_len(nn)));
                } else if (fn == PTG_FN_LP_CL_CEILING) {
                    if (!(x[0] >= 1.0 && x[0] <= 2.0e5) || double(int(x[0])) != x[0] || !(x[1] > 0.0))
                        return fail(c, PTG_E_ARG, "LP classical ceiling probe needs an integer 1 <= n <= 2e5 and a > 0");
                } else if (!(x[3] >= 0.0 && x[3] <= 2.0e5)) {
                    return fail(c, PTG_E_ARG, "LP sample probe needs 0 <= fixed n <= 2e5");
                }
            }
        }
        if (n * len * 5 * 8 > (size_t(16) << 30)) return fail(c, PTG_E_ARG, "probe batch too large for its J(n, x, y) workspace");
        double *din = nullptr, *dout = nullptr, *dscr = nullptr;
        HIPCHK(c, hipMalloc(&din, n * in_w * 8));
        HIPCHK(c, hipMalloc(&dout, n * out_w * 8));
        if (len) HIPCHK(c, hipMalloc(&dscr, n * len * 5 * 8));
        HIPCHK(c, scopy(c, din, in, n * in_w * 8, hipMemcpyHostToDevice));
        HIPCHK(c, sfill(c, dout, 0, n * out_w * 8));
        LmaArgs la{};
        la.tables = c->d_lma;
        la.sc = c->lsc;
        la.lp_ceil = c->d_lp_ceil;
        la.bw = c->d_bw;
        la.bsc = c->bsc;
        la.cp = c->d_cp;
        la.csc = c->csc;
        la.lcfa = c->d_lcfa;
        la.scratch = dscr;
        launch_eval_lma(fn, din, dout, uint32_t(n), uint32_t(in_w), uint32_t(out_w), la, key_from_seed(0), c->stream);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, scopy(c, out, dout, n * out_w * 8, hipMemcpyDeviceToHost));
        (void)hipFree(din); (void)hipFree(dout);
        if (dscr) (void)hipFree(dscr);
        return PTG_OK;
    }
    HIPCHK(c, hipSetDevice(c->device));
    double *din = nullptr, *dout = nullptr;
    HIPCHK(c, hipMalloc(&din, n * in_w * 8));
    HIPCHK(c, hipMalloc(&dout, n * out_w * 8));
    HIPCHK(c, scopy(c, din, in, n * in_w * 8, hipMemcpyHostToDevice));
    HIPCHK(c, sfill(c, dout, 0, n * out_w * 8));
    EvalArgs e{};
    e.fn = fn;
    e.in = din;
    e.out = dout;
    e.n = uint32_t(n);
    e.in_w = uint32_t(in_w);
    e.out_w = uint32_t(out_w);
    e.lcfa = c->d_lcfa;
    e.sc = c->sc;
    e.laser = c->laser;
    launch_eval(e, c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, scopy(c, out, dout, n * out_w * 8, hipMemcpyDeviceToHost));
    (void)hipFree(din);
    (void)hipFree(dout);
    return PTG_OK;
}

int ptg_set_batch(ptg_ctx* c, const ptg_laser_desc* lasers, const ptg_beam_desc* beams, int n_coll, uint64_t seed) {
    if (!c || !lasers || !beams) return fail(c, PTG_E_ARG, "null argument");
    if (n_coll < 1 || n_coll > 256) return fail(c, PTG_E_ARG, "a batch holds 1 to 256 collisions");
    const ptg_laser_desc& d0 = lasers[0];
    if (d0.kind != PTG_PLANE_WAVE && d0.kind != PTG_FOCUSED_LASER)
        return fail(c, PTG_E_UNSUPPORTED, "a batch of collisions runs the LMA fields (PlaneWave, FocusedLaser)");
    for (int k = 1; k < n_coll; ++k) {
        const ptg_laser_desc& d = lasers[k];
        if (d.kind != d0.kind || d.envelope != d0.envelope || d.polarization != d0.polarization ||
            d.finite_bandwidth != d0.finite_bandwidth || d.wavelength != d0.wavelength || d.waist != d0.waist ||
            d.n_cycles != d0.n_cycles || d.pol_angle != d0.pol_angle || d.chirp_b != d0.chirp_b)
            return fail(c, PTG_E_ARG, "the lasers of a batch may differ only in a0");
    }
    uint64_t total = 0;
    for (int k = 0; k < n_coll; ++k) {
        const ptg_beam_desc& b = beams[k];
        if (b.n == 0) return fail(c, PTG_E_ARG, "every beam of a batch needs primaries");
        if (b.species < 0 || b.species > 2 || b.radial < 0 || b.radial > 2) return fail(c, PTG_E_ARG, "bad species or radial distribution");
        if (!(std::fabs(b.energy_chirp) <= 1.0)) return fail(c, PTG_E_ARG, "|energy_chirp| must be <= 1");
        total += b.n;
    }
    if (total > 0xffffffffull) return fail(c, PTG_E_ARG, "a batch holds fewer than 2^32 primaries");
    std::vector<LaserDev> L(n_coll);
    double a0_max = 0.0;
    for (int k = 0; k < n_coll; ++k) {
        int rc = make_laser(c, &lasers[k], L[k]);
        if (rc) return rc;
        a0_max = std::max(a0_max, lasers[k].a0);
    }
    int rc = ptg_set_laser(c, &d0);  // the batch's common field parameters (time step, polarization, bandwidth)
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    if (!c->d_lasers) HIPCHK(c, hipMalloc(&c->d_lasers, 256 * sizeof(LaserDev)));
    HIPCHK(c, scopy(c, c->d_lasers, L.data(), n_coll * sizeof(LaserDev), hipMemcpyHostToDevice));
    if (c->prim.cap < total) HIPCHK(c, alloc_buf(c->prim, total));
    rc = ensure_n0(c, total);
    if (rc) return rc;
    if (c->coll_cap < total) {
        if (c->d_coll) (void)hipFree(c->d_coll);
        c->d_coll = nullptr;
        c->coll_cap = 0;
        HIPCHK(c, hipMalloc(&c->d_coll, total));
        c->coll_cap = total;
    }
    std::vector<uint8_t> coll(total);
    uint64_t off = 0;
    for (int k = 0; k < n_coll; ++k) {
        const ptg_beam_desc& b = beams[k];
        ptg_laser_desc dk = lasers[k];
        const double z0 = ideal_initial_z(&dk, L[k]) + 3.0 * b.length;  // the driver's initial_z (src/main.rs:1062-1066)
        BeamDev bd{};
        set_spectrum(c, bd);
        bd.n = uint32_t(b.n);
        bd.species = b.species;
        bd.radial = b.radial;
        bd.gamma = b.gamma;
        bd.sigma = b.sigma;
        bd.rho = -b.energy_chirp;
        bd.radius = b.radius;
        bd.max_radius = b.max_radius;
        bd.length = b.length;
        bd.rms_div = b.rms_div;
        bd.angle = b.angle;
        bd.plane_angle = b.collision_plane_angle;
        bd.off_x = b.offset[0]; bd.off_y = b.offset[1]; bd.off_z = b.offset[2];
        for (int q = 0; q < 4; ++q) bd.pol[q] = b.pol[q];
        bd.weight = b.weight;
        bd.initial_z = z0;
        // collision 0 draws as ptg_build_beam with the same seed would
        bd.key = mix64(key_from_seed(seed) ^ 0x6265616d6265616dULL ^ (0x9e3779b97f4a7c15ULL * uint64_t(k)));
        bd.off = uint32_t(off);
        bd.n_total = uint32_t(total);
        launch_build_beam(bd, c->prim.v, c->d_n0, c->stream);
        HIPCHK(c, hipGetLastError());
        std::fill(coll.begin() + off, coll.begin() + off + b.n, uint8_t(k));
        off += b.n;
    }
    HIPCHK(c, scopy(c, c->d_coll, coll.data(), total, hipMemcpyHostToDevice));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->n_prim = total;
    c->n_coll = n_coll;
    c->batch_a0_max = a0_max;
    return PTG_OK;
}

int ptg_set_launch_mode(ptg_ctx* c, int mode) {
    if (!c) return fail(c, PTG_E_ARG, "null argument");
    if (mode != PTG_LAUNCH_EAGER && mode != PTG_LAUNCH_GRAPH) return fail(c, PTG_E_ARG, "unknown launch mode");
    c->launch_mode = mode;
    return PTG_OK;
}

int ptg_autotune(ptg_ctx* c, const ptg_collide_options* o, uint64_t seed, uint64_t max_primaries, ptg_tune_result* out) {
    if (!c || !o) return PTG_E_ARG;
    if (!(c->laser.kind == K_PLANE_WAVE || c->laser.kind == K_FOCUSED_LASER))
        return fail(c, PTG_E_UNSUPPORTED, "the autotuner tunes the LMA lepton step kernel");
    if (c->n_prim == 0) return fail(c, PTG_E_STATE, "no primaries loaded");
    const size_t keep = c->n_prim;
    c->n_prim = std::min<size_t>(keep, max_primaries ? max_primaries : keep);
    const int blocks[3] = {64, 128, 256}, grids[3] = {2, 4, 8};
    ptg_tune_result r{};
    double best = 1e300;
    uint64_t emissions0 = 0;
    r.consistent = 1;
    int rc = PTG_OK;
    for (int bi = 0; bi < 3 && !rc; ++bi)
        for (int gi = 0; gi < 3 && !rc; ++gi) {
            c->step_block = blocks[bi];
            c->step_grid_per_cu = grids[gi];
            ptg_counters cnt{};
            rc = ptg_run(c, o, seed, &cnt);
            if (rc) break;
            const int k = r.n_candidates++;
            r.cand_block[k] = blocks[bi];
            r.cand_grid_per_cu[k] = grids[gi];
            r.cand_ms[k] = cnt.lepton_ms;
            if (k == 0) emissions0 = cnt.emissions;
            else if (cnt.emissions != emissions0) r.consistent = 0;
            if (cnt.lepton_ms < best) {
                best = cnt.lepton_ms;
                r.block = blocks[bi];
                r.grid_per_cu = grids[gi];
            }
        }
    c->n_prim = keep;
    if (rc) {
        c->step_block = 256;
        c->step_grid_per_cu = 4;
        return rc;
    }
    c->step_block = r.block;
    c->step_grid_per_cu = r.grid_per_cu;
    if (out) *out = r;
    return PTG_OK;
}

static int check_surrogate_box(ptg_ctx* c, double a_min, double a_max, double eta_min, double eta_max) {
    if (!(a_min > 0.0 && a_max > a_min && a_max <= 20.0))
        return fail(c, PTG_E_ARG, "surrogate domain needs 0 < a_min < a_max <= 20 (the LMA rate tables end at a = 20)");
    if (!(eta_min > 0.0 && eta_max > eta_min && eta_max <= 2.0))
        return fail(c, PTG_E_ARG, "surrogate domain needs 0 < eta_min < eta_max <= 2");
    return PTG_OK;
}

int ptg_surrogate_load(ptg_ctx* c, const ptg_surrogate_model* model) {
    if (!c || !model) return PTG_E_ARG;
    if (model->polarization != PTG_LINEAR && model->polarization != PTG_CIRCULAR) return fail(c, PTG_E_ARG, "unknown polarization");
    int rc = check_surrogate_box(c, model->a_min, model->a_max, model->eta_min, model->eta_max);
    if (rc) return rc;
    if (!(model->a_use_min >= 0.0 && std::isfinite(model->a_use_min) && model->eta_use_min >= 0.0 &&
          std::isfinite(model->eta_use_min)))
        return fail(c, PTG_E_ARG, "surrogate a_use_min and eta_use_min must be finite and >= 0");
    for (int i = 0; i < PTG_MLP_NPARAM; ++i)
        if (!std::isfinite(model->params[i])) return fail(c, PTG_E_ARG, "surrogate parameters must be finite");
    HIPCHK(c, hipSetDevice(c->device));
    const int p = model->polarization == PTG_CIRCULAR ? 1 : 0;
    HIPCHK(c, c->surr[p].init_params(model->params, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->surr_box[p][0] = model->a_min;
    c->surr_box[p][1] = model->a_max;
    c->surr_box[p][2] = model->eta_min;
    c->surr_box[p][3] = model->eta_max;
    c->surr_box[p][4] = model->a_use_min;
    c->surr_box[p][5] = model->eta_use_min;
    c->surr_loaded[p] = true;
    c->surr_valid[p] = model->valid != 0;
    return PTG_OK;
}

int ptg_mlp_probe(ptg_ctx* c, const float* params, const float* x, const float* t, size_t n, float w_t, float* out, float* grads,
                  float* loss) {
    if (!c || !params || !x || !t || !out || !grads || !loss) return PTG_E_ARG;
    if (n == 0 || n % 32 != 0 || n > (size_t(1) << 20)) return fail(c, PTG_E_ARG, "the probe needs n rows, a multiple of 32, at most 2^20");
    HIPCHK(c, hipSetDevice(c->device));
    MlpNet net;
    int rc = PTG_OK;
    hipError_t e = net.init_params(params, c->stream);
    if (!e) e = net.reserve(int(n));
    if (!e) e = hipMemcpyAsync(net.X, x, n * 8 * sizeof(float), hipMemcpyHostToDevice, c->stream);
    if (!e) e = hipMemcpyAsync(net.T, t, n * 8 * sizeof(float), hipMemcpyHostToDevice, c->stream);
    if (!e) {
        net.forward(int(n), true, c->stream);
        net.backward(int(n), w_t, c->stream);
        e = hipGetLastError();
    }
    if (!e) e = hipMemcpyAsync(out, net.O, n * mlp::OUT * sizeof(float), hipMemcpyDeviceToHost, c->stream);
    if (!e) e = hipMemcpyAsync(grads, net.G, mlp::NPARAM * sizeof(float), hipMemcpyDeviceToHost, c->stream);
    if (!e) e = hipMemcpyAsync(loss, net.loss, 3 * sizeof(float), hipMemcpyDeviceToHost, c->stream);
    if (!e) e = hipStreamSynchronize(c->stream);
    if (e) rc = fail(c, PTG_E_HIP, std::string("mlp probe: ") + hipGetErrorString(e));
    net.release();
    return rc;
}

int ptg_surrogate_train(ptg_ctx* c, const ptg_surrogate_train_desc* d, ptg_surrogate_model* model_out, ptg_surrogate_report* rep) {
    if (!c || !d) return PTG_E_ARG;
    if (d->polarization != PTG_LINEAR && d->polarization != PTG_CIRCULAR) return fail(c, PTG_E_ARG, "unknown polarization");
    if (!(d->groups >= 64 && d->groups <= (1 << 22))) return fail(c, PTG_E_ARG, "groups must be in [64, 2^22]");
    if (!(d->steps >= 1 && d->steps <= 10000000)) return fail(c, PTG_E_ARG, "steps must be in [1, 1e7]");
    if (!(d->batch >= 64 && d->batch % 64 == 0 && d->batch <= (1 << 16))) return fail(c, PTG_E_ARG, "batch must be a multiple of 64 in [64, 65536]");
    if (!(d->lr > 0.0 && d->lr < 1.0)) return fail(c, PTG_E_ARG, "lr must be in (0, 1)");
    int rc = check_surrogate_box(c, d->a_min, d->a_max, d->eta_min, d->eta_max);
    if (rc) return rc;
    const bool circ = d->polarization == PTG_CIRCULAR;
    if (circ && !c->d_cp) return fail(c, PTG_E_UNSUPPORTED, "CP emission needs the regenerated CP CDF table");
    HIPCHK(c, hipSetDevice(c->device));
    LmaArgs m{};
    m.tables = c->d_lma;
    m.sc = c->lsc;
    m.lp_ceil = c->d_lp_ceil;
    m.cp = c->d_cp;
    m.csc = c->csc;
    m.lcfa = c->d_lcfa;
    m.mode = circ ? LMA_CP : LMA_LP;
    const double box[6] = {d->a_min, d->a_max, d->eta_min, d->eta_max, 0.0, 0.0};
    SurrArgs S{};
    surr_box_args(box, S);
    S.key = key_from_seed(d->seed);
    S.n_groups = uint32_t(d->groups);
    // ---- exact-sampler groups (one wave per group, up to 4096 waves)
    const size_t lanes = std::min<size_t>(size_t(d->groups) * 64, 262144);
    double* gh = nullptr;
    float* gs = nullptr;
    float* drows = nullptr;
    MlpNet net;
    auto cleanup = [&]() {
        if (gh) (void)hipFree(gh);
        if (gs) (void)hipFree(gs);
        if (drows) (void)hipFree(drows);
        net.release();
    };
    auto hip_fail = [&](hipError_t e) {
        cleanup();
        return fail(c, e == hipErrorOutOfMemory ? PTG_E_OOM : PTG_E_HIP, std::string("surrogate training: ") + hipGetErrorString(e));
    };
    hipError_t e;
    if ((e = hipMalloc(&gh, size_t(d->groups) * 4 * sizeof(double)))) return hip_fail(e);
    if ((e = hipMalloc(&gs, size_t(d->groups) * 128 * sizeof(float)))) return hip_fail(e);
    S.gh = gh;
    S.gs = gs;
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    launch_surr_groups(m, S, int(lanes), c->stream);
    if ((e = hipGetLastError())) return hip_fail(e);
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    std::vector<double> hgh(size_t(d->groups) * 4);
    std::vector<float> hgs(size_t(d->groups) * 128);
    if ((e = hipMemcpyAsync(hgh.data(), gh, hgh.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream))) return hip_fail(e);
    if ((e = hipMemcpyAsync(hgs.data(), gs, hgs.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream))) return hip_fail(e);
    if ((e = hipStreamSynchronize(c->stream))) return hip_fail(e);
    float gen_ms = 0.0f;
    (void)hipEventElapsedTime(&gen_ms, c->ev0, c->ev1);
    // rows: each group's 64 samples sorted by v, the k-th at the level (k + 1) / 65
    // (the mean of the k-th of 64 ordered uniforms, so that the mean-square fit of
    // the ordered samples is the quantile function, tails included); every eighth
    // group held out for validation
    std::vector<float> rows, vrows;
    int used = 0;
    for (int g = 0; g < d->groups; ++g) {
        const double* h = &hgh[size_t(g) * 4];
        if (!(h[3] > 0.0)) continue;
        ++used;
        const float* smp = &hgs[size_t(g) * 128];
        int idx[64];
        for (int k = 0; k < 64; ++k) idx[k] = k;
        std::sort(idx, idx + 64, [&](int x, int y) { return smp[2 * x] < smp[2 * y]; });
        std::vector<float>& dst = (g % 8 == 7) ? vrows : rows;
        for (int k = 0; k < 64; ++k) {
            const float row[8] = {float(h[0]), float(h[1]), float(h[2]), float((k + 1.0) / 65.0),
                                  float(std::log(std::max(h[3], 0.9)) / 5.0), smp[2 * idx[k]], smp[2 * idx[k] + 1], 0.0f};
            dst.insert(dst.end(), row, row + 8);
        }
    }
    if (rows.empty() || vrows.empty()) {
        cleanup();
        return fail(c, PTG_E_STATE, "surrogate training: the exact sampler produced too few groups");
    }
    if ((e = hipMalloc(&drows, rows.size() * sizeof(float)))) return hip_fail(e);
    if ((e = hipMemcpyAsync(drows, rows.data(), rows.size() * sizeof(float), hipMemcpyHostToDevice, c->stream))) return hip_fail(e);
    // ---- training
    std::vector<float> p0;
    mlp_init_params(p0, d->seed ^ 0x9e3779b97f4a7c15ull);
    if ((e = net.init_params(p0.data(), c->stream))) return hip_fail(e);
    const size_t nv = std::min<size_t>(vrows.size() / 8, 65536);  // held-out rows validated (whole groups)
    const int vM = int((nv + 31) / 32 * 32);
    if ((e = net.reserve(std::max(d->batch, vM)))) return hip_fail(e);
    S.rows = drows;
    S.n_sampler_rows = uint32_t(rows.size() / 8);
    S.tx = net.X;
    S.tt = net.T;
    S.batch = d->batch;
    S.n_table_rows = d->batch / 2;
    double lsum[3] = {0.0, 0.0, 0.0};
    int lcount = 0;
    const int tail = std::max(1, d->steps / 10);
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    for (int step = 0; step < d->steps; ++step) {
        S.step = uint32_t(step);
        launch_surr_batch(m, S, c->stream);
        net.forward(d->batch, true, c->stream);
        net.backward(d->batch, 1.0f, c->stream);
        const double lr = d->lr * 0.5 * (1.0 + std::cos(PI_H * double(step) / double(d->steps)));
        net.adam(float(lr), c->stream);
        if (step >= d->steps - tail) {
            float l[3];
            if ((e = hipMemcpyAsync(l, net.loss, sizeof(l), hipMemcpyDeviceToHost, c->stream))) return hip_fail(e);
            if ((e = hipStreamSynchronize(c->stream))) return hip_fail(e);
            lsum[0] += l[0] / d->batch;  // harmonic: every row
            lsum[1] += l[1] / (d->batch / 2);  // fraction, angle: the sampler rows
            lsum[2] += l[2] / (d->batch / 2);
            ++lcount;
        }
    }
    if ((e = hipGetLastError())) return hip_fail(e);
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    if ((e = hipStreamSynchronize(c->stream))) return hip_fail(e);
    float train_ms = 0.0f;
    (void)hipEventElapsedTime(&train_ms, c->ev0, c->ev1);
    // ---- validation on the held-out groups
    std::vector<float> vx(size_t(vM) * 8, 0.0f), vo(size_t(vM) * mlp::OUT);
    for (size_t r = 0; r < nv; ++r) {
        const float* row = &vrows[r * 8];
        vx[r * 8 + 0] = float((row[0] - S.la_c) / S.la_h);
        vx[r * 8 + 1] = float((row[1] - S.le_c) / S.le_h);
        vx[r * 8 + 2] = 2.0f * row[2] - 1.0f;
        vx[r * 8 + 3] = 2.0f * row[3] - 1.0f;
        vx[r * 8 + 4] = 0.0f;
    }
    auto infer = [&]() -> hipError_t {
        hipError_t err = hipMemcpyAsync(net.X, vx.data(), vx.size() * sizeof(float), hipMemcpyHostToDevice, c->stream);
        if (err) return err;
        net.forward(vM, false, c->stream);
        if ((err = hipMemcpyAsync(vo.data(), net.O, vo.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream))) return err;
        return hipStreamSynchronize(c->stream);
    };
    if ((e = infer())) return hip_fail(e);
    double h_rel = 0.0, v_abs = 0.0;
    for (size_t r = 0; r < nv; ++r) {
        const double nu = std::exp(5.0 * double(vrows[r * 8 + 4]));
        const double n_true = std::max(1.0, std::ceil(nu));
        const double n_pred = std::max(1.0, std::ceil(std::exp(5.0 * double(vo[r * mlp::OUT]))));
        h_rel += std::fabs(n_pred - n_true) / n_true;
        v_abs += std::fabs(std::min(std::max(double(vo[r * mlp::OUT + 1]), 0.0), 1.0) - double(vrows[r * 8 + 5]));
    }
    h_rel /= double(nv);
    v_abs /= double(nv);
    // probability integral transforms of the held-out samples under the network:
    // the level u2 at which the network's v reaches the sample's v (bisection),
    // then, at that level, the u3 at which its theta reaches the sample's theta;
    // uniform if the network has the samples' distribution
    // v as the inference kernel forms it: the network at the level clamped to
    // [1/65, 64/65], continued as sqrt(u) towards 0 and linearly towards 1 beyond
    auto level_in = [](int col, double u) { return col == 3 ? std::min(std::max(u, 1.0 / 65.0), 64.0 / 65.0) : u; };
    auto value_out = [](int col, double u, double o) {
        if (col != 1) return o;
        const double v = std::min(std::max(o, 0.0), 1.0);
        if (u < 1.0 / 65.0) return v * std::sqrt(u * 65.0);
        if (u > 64.0 / 65.0) return 1.0 - (1.0 - v) * (1.0 - u) * 65.0;
        return v;
    };
    auto pit = [&](int col_in, int col_out, int col_target, std::vector<double>& level) -> hipError_t {
        std::vector<double> lo(nv, 0.0), hi(nv, 1.0);
        for (int it = 0; it < 20; ++it) {
            for (size_t r = 0; r < nv; ++r) vx[r * 8 + col_in] = float(2.0 * level_in(col_in, 0.5 * (lo[r] + hi[r])) - 1.0);
            hipError_t err = infer();
            if (err) return err;
            for (size_t r = 0; r < nv; ++r) {
                const double mid = 0.5 * (lo[r] + hi[r]);
                if (value_out(col_out, mid, vo[r * mlp::OUT + col_out]) < vrows[r * 8 + col_target]) lo[r] = mid;
                else hi[r] = mid;
            }
        }
        level.resize(nv);
        for (size_t r = 0; r < nv; ++r) level[r] = 0.5 * (lo[r] + hi[r]);
        return hipSuccess;
    };
    auto ks_uniform = [&](std::vector<double> u) {
        std::sort(u.begin(), u.end());
        double dmax = 0.0;
        const double N = double(u.size());
        for (size_t i = 0; i < u.size(); ++i) dmax = std::max(dmax, std::max(double(i + 1) / N - u[i], u[i] - double(i) / N));
        return dmax;
    };
    std::vector<double> pit_v, pit_t;
    for (size_t r = 0; r < nv; ++r) vx[r * 8 + 4] = 0.0f;
    if ((e = pit(3, 1, 5, pit_v))) return hip_fail(e);
    const double ks_v = ks_uniform(pit_v);
    double ks_t = 0.0;
    if (!circ) {
        for (size_t r = 0; r < nv; ++r) vx[r * 8 + 3] = float(2.0 * level_in(3, pit_v[r]) - 1.0);
        if ((e = pit(4, 2, 6, pit_t))) return hip_fail(e);
        ks_t = ks_uniform(pit_t);
    }
    const bool valid = h_rel < 0.05 && ks_v < 0.02 && ks_t < 0.02;
    // ---- export and load
    ptg_surrogate_model* mdl = new ptg_surrogate_model();
    mdl->polarization = d->polarization;
    mdl->valid = valid ? 1 : 0;
    mdl->a_min = d->a_min;
    mdl->a_max = d->a_max;
    mdl->eta_min = d->eta_min;
    mdl->eta_max = d->eta_max;
    if ((e = net.download_params(mdl->params, c->stream))) {
        delete mdl;
        return hip_fail(e);
    }
    cleanup();
    rc = ptg_surrogate_load(c, mdl);
    if (!rc && model_out) std::memcpy(model_out, mdl, sizeof(*mdl));
    delete mdl;
    if (rc) return rc;
    if (rep) {
        std::memset(rep, 0, sizeof(*rep));
        for (int k = 0; k < 3; ++k) rep->loss[k] = lcount ? lsum[k] / lcount : 0.0;
        rep->val_harmonic_rel = h_rel;
        rep->val_fraction_abs = v_abs;
        rep->ks_fraction = ks_v;
        rep->ks_angle = ks_t;
        rep->valid = valid ? 1 : 0;
        rep->groups_used = used;
        rep->table_rows = uint64_t(d->steps) * uint64_t(d->batch / 2);
        rep->sampler_rows = uint64_t(d->steps) * uint64_t(d->batch - d->batch / 2);
        rep->gen_ms = gen_ms;
        rep->train_ms = train_ms;
    }
    return PTG_OK;
}


}  // extern "C"
