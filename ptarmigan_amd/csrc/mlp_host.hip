// Host driver of the surrogate's multilayer perceptron.
#include "mlp_host.h"

#include <cmath>
#include <random>

namespace ptg {

void mlp_init_params(std::vector<float>& p, uint64_t seed) {
    using namespace mlp;
    p.assign(NPARAM, 0.0f);
    std::mt19937_64 g(seed);
    auto fill = [&](int off, int rows, int cols, int fan_in) {
        const float lim = std::sqrt(6.0f / float(fan_in));
        std::uniform_real_distribution<float> d(-lim, lim);
        for (int r = 0; r < rows; ++r)
            for (int c = 0; c < cols; ++c) p[off + r * cols + c] = d(g);
    };
    fill(W1, H, IN, 5);
    for (int r = 0; r < H; ++r)
        for (int c = 5; c < IN; ++c) p[W1 + r * IN + c] = 0.0f;  // padding inputs
    fill(W2, H, H, H);
    fill(W3, H, H, H);
    fill(W4, 3, H, H);  // three outputs; rows 3..OUT-1 stay zero
    for (int c = 0; c < H; ++c) p[W4 + c] *= 0.1f, p[W4 + H + c] *= 0.1f, p[W4 + 2 * H + c] *= 0.1f;
}

hipError_t MlpNet::init_params(const float* host_params, hipStream_t s) {
    using namespace mlp;
    hipError_t e;
    if (!P) {
        if ((e = hipMalloc(&P, NPARAM * sizeof(float)))) return e;
        if ((e = hipMalloc(&G, NPARAM * sizeof(float)))) return e;
        if ((e = hipMalloc(&m1, NPARAM * sizeof(float)))) return e;
        if ((e = hipMalloc(&m2, NPARAM * sizeof(float)))) return e;
        if ((e = hipMalloc(&Wb, NBF * sizeof(uint16_t)))) return e;
        if ((e = hipMalloc(&loss, 4 * sizeof(float)))) return e;
    }
    if ((e = hipMemcpyAsync(P, host_params, NPARAM * sizeof(float), hipMemcpyHostToDevice, s))) return e;
    if ((e = hipMemsetAsync(m1, 0, NPARAM * sizeof(float), s))) return e;
    if ((e = hipMemsetAsync(m2, 0, NPARAM * sizeof(float), s))) return e;
    if ((e = hipMemsetAsync(G, 0, NPARAM * sizeof(float), s))) return e;
    step = 0;
    launch_publish(P, Wb, s);
    return hipGetLastError();
}

hipError_t MlpNet::reserve(int rows) {
    using namespace mlp;
    if (rows <= cap) return hipSuccess;
    auto f = [&](float*& p, size_t n) {
        if (p) (void)hipFree(p);
        p = nullptr;
        return hipMalloc(&p, n * sizeof(float));
    };
    auto b = [&](uint16_t*& p, size_t n) {
        if (p) (void)hipFree(p);
        p = nullptr;
        return hipMalloc(&p, n * sizeof(uint16_t));
    };
    const size_t R = size_t(rows);
    hipError_t e;
    if ((e = f(X, R * IN)) || (e = f(T, R * 8)) || (e = f(Z1, R * H)) || (e = f(Z2, R * H)) || (e = f(Z3, R * H)) ||
        (e = f(O, R * OUT)) || (e = f(Df, R * H)) || (e = f(dOf, R * OUT)))
        return e;
    if ((e = b(A1, R * H)) || (e = b(A2, R * H)) || (e = b(A3, R * H)) || (e = b(A1t, R * H)) || (e = b(A2t, R * H)) ||
        (e = b(A3t, R * H)) || (e = b(D2, R * H)) || (e = b(D3, R * H)) || (e = b(D2t, R * H)) || (e = b(D3t, R * H)) ||
        (e = b(dO, R * OUT)) || (e = b(dOt, R * OUT)))
        return e;
    cap = rows;
    return hipSuccess;
}

void MlpNet::forward(int M, bool train, hipStream_t s, const unsigned long long* m_dev) {
    using namespace mlp;
    launch_layer1_forward(X, P, M, A1, train ? A1t : nullptr, train ? Z1 : nullptr, s, m_dev);
    GemmArgs g{};
    // layer 2: A2 = SiLU(A1 W2^T + b2)
    g = GemmArgs{A1, H, Wb + BW2, H, M, H, H, P + B2, 1, nullptr, 0, train ? Z2 : nullptr, H, A2, H, train ? A2t : nullptr, M, H,
                 m_dev};
    launch_gemm(g, s);
    // layer 3
    g = GemmArgs{A2, H, Wb + BW3, H, M, H, H, P + B3, 1, nullptr, 0, train ? Z3 : nullptr, H, A3, H, train ? A3t : nullptr, M, H,
                 m_dev};
    launch_gemm(g, s);
    // layer 4 (linear): O = A3 W4^T + b4
    g = GemmArgs{A3, H, Wb + BW4, H, M, OUT, H, P + B4, 0, nullptr, 0, O, OUT, nullptr, 0, nullptr, 0, OUT, m_dev};
    launch_gemm(g, s);
}

void MlpNet::backward(int M, float w_t, hipStream_t s) {
    using namespace mlp;
    (void)hipMemsetAsync(loss, 0, 4 * sizeof(float), s);
    launch_loss(O, T, M, w_t, dO, dOt, dOf, loss, s);
    GemmArgs g{};
    // dW4 = dO^T A3, db4 = colsum dO
    g = GemmArgs{dOt, M, A3t, M, OUT, H, M, nullptr, 0, nullptr, 0, G + W4, H, nullptr, 0, nullptr, 0, H};
    launch_gemm(g, s);
    launch_colsum(dOf, M, OUT, OUT, G + B4, s);
    // d3 = (dO W4) * SiLU'(z3)
    g = GemmArgs{dO, OUT, Wb + BW4T, OUT, M, H, OUT, nullptr, 2, Z3, H, Df, H, D3, H, D3t, M, H};
    launch_gemm(g, s);
    launch_colsum(Df, M, H, H, G + B3, s);
    // dW3 = d3^T A2
    g = GemmArgs{D3t, M, A2t, M, H, H, M, nullptr, 0, nullptr, 0, G + W3, H, nullptr, 0, nullptr, 0, H};
    launch_gemm(g, s);
    // d2 = (d3 W3) * SiLU'(z2)
    g = GemmArgs{D3, H, Wb + BW3T, H, M, H, H, nullptr, 2, Z2, H, Df, H, D2, H, D2t, M, H};
    launch_gemm(g, s);
    launch_colsum(Df, M, H, H, G + B2, s);
    // dW2 = d2^T A1
    g = GemmArgs{D2t, M, A1t, M, H, H, M, nullptr, 0, nullptr, 0, G + W2, H, nullptr, 0, nullptr, 0, H};
    launch_gemm(g, s);
    // d1 = (d2 W2) * SiLU'(z1), FP32 only (layer 1 runs in FP32)
    g = GemmArgs{D2, H, Wb + BW2T, H, M, H, H, nullptr, 2, Z1, H, Df, H, nullptr, 0, nullptr, 0, H};
    launch_gemm(g, s);
    launch_colsum(Df, M, H, H, G + B1, s);
    launch_layer1_wgrad(Df, X, M, G, s);
}

void MlpNet::adam(float lr, hipStream_t s) {
    launch_adam(P, G, m1, m2, ++step, lr, s);
    launch_publish(P, Wb, s);
}

hipError_t MlpNet::download_params(float* host_params, hipStream_t s) const {
    hipError_t e = hipMemcpyAsync(host_params, P, mlp::NPARAM * sizeof(float), hipMemcpyDeviceToHost, s);
    if (e) return e;
    return hipStreamSynchronize(s);
}

void MlpNet::release() {
    for (float** p : {&P, &G, &m1, &m2, &loss, &X, &T, &Z1, &Z2, &Z3, &O, &Df, &dOf})
        if (*p) (void)hipFree(*p), *p = nullptr;
    for (uint16_t** p : {&Wb, &A1, &A2, &A3, &A1t, &A2t, &A3t, &D2, &D3, &D2t, &D3t, &dO, &dOt})
        if (*p) (void)hipFree(*p), *p = nullptr;
    cap = 0;
}

}  // namespace ptg
