// Host driver of the surrogate's multilayer perceptron (kernels in kern_mlp.hip):
// device buffers, the forward pass (inference or training), the backward pass
// and the Adam step.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

#include "kern_mlp.h"

namespace ptg {

struct MlpNet {
    float* P = nullptr;        // FP32 parameters (mlp:: layout)
    float* G = nullptr;        // gradients
    float* m1 = nullptr;       // Adam moments
    float* m2 = nullptr;
    uint16_t* Wb = nullptr;    // BF16 copies
    float* loss = nullptr;     // [3] loss sums of the last batch
    int step = 0;              // Adam steps taken
    // activations for up to `cap` rows
    int cap = 0;
    float *X = nullptr, *T = nullptr, *Z1 = nullptr, *Z2 = nullptr, *Z3 = nullptr, *O = nullptr, *Df = nullptr, *dOf = nullptr;
    uint16_t *A1 = nullptr, *A2 = nullptr, *A3 = nullptr, *A1t = nullptr, *A2t = nullptr, *A3t = nullptr;
    uint16_t *D2 = nullptr, *D3 = nullptr, *D2t = nullptr, *D3t = nullptr, *dO = nullptr, *dOt = nullptr;

    hipError_t init_params(const float* host_params, hipStream_t s);  // upload + publish; allocates P..Wb
    hipError_t reserve(int rows);                                     // activation buffers for `rows` (multiple of 32)
    // Forward over M rows of X (M a multiple of 32): O [M][OUT]. `train` keeps
    // what the backward pass needs. With m_dev set, M bounds the rows and only
    // the first *m_dev (rounded up to 32) are computed.
    void forward(int M, bool train, hipStream_t s, const unsigned long long* m_dev = nullptr);
    // Loss of T against O, gradients into G (all M rows), then no update.
    void backward(int M, float w_t, hipStream_t s);
    void adam(float lr, hipStream_t s);
    hipError_t download_params(float* host_params, hipStream_t s) const;
    void release();
};

// Kaiming-uniform initial parameters, unused output rows zero.
void mlp_init_params(std::vector<float>& p, uint64_t seed);

}  // namespace ptg
