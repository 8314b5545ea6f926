// The surrogate's multilayer perceptron on the device: 5 inputs -> 128 -> 128
// -> 128 -> 3 outputs, SiLU activations. Layer 1 (5 inputs) runs in FP32 on the
// vector ALUs so that the uniform draws among the inputs keep full resolution;
// layers 2-4 run as BF16 MFMA GEMMs with FP32 accumulation and fused bias /
// SiLU epilogues. Training keeps FP32 master weights (Adam) and publishes BF16
// copies (and their transposes for the backward pass).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace ptg {

namespace mlp {
constexpr int IN = 8;      // padded input width (5 used)
constexpr int H = 128;     // hidden width
constexpr int OUT = 32;    // padded output width (3 used)
// flat FP32 parameter layout
constexpr int W1 = 0;                  // [H][IN]
constexpr int B1 = W1 + H * IN;        // [H]
constexpr int W2 = B1 + H;             // [H][H]
constexpr int B2 = W2 + H * H;
constexpr int W3 = B2 + H;             // [H][H]
constexpr int B3 = W3 + H * H;
constexpr int W4 = B3 + H;             // [OUT][H]
constexpr int B4 = W4 + OUT * H;       // [OUT]
constexpr int NPARAM = B4 + OUT;
// BF16 copies (uint16 elements)
constexpr int BW2 = 0, BW2T = BW2 + H * H, BW3 = BW2T + H * H, BW3T = BW3 + H * H, BW4 = BW3T + H * H,
              BW4T = BW4 + OUT * H, BW1 = BW4T + OUT * H, NBF = BW1 + H * IN;
}  // namespace mlp

// C[M][N] (+ epilogue) = A[M][K] * Bt[N][K]^T, BF16 operands, FP32 accumulation.
// M, N multiples of 32, K a multiple of 8, leading dimensions multiples of 4.
struct GemmArgs {
    const uint16_t* A; int lda;
    const uint16_t* Bt; int ldb;
    int M, N, K;
    const float* bias;           // [N] or null
    int act;                     // 0: none, 1: SiLU, 2: times SiLU'(zin)
    const float* zin; int ldzi;  // act 2
    float* z; int ldz;           // FP32 pre-activation (act 1) or result (acts 0, 2), or null
    uint16_t* C; int ldc;        // BF16 result (after activation), or null
    uint16_t* Ct; int ldct;      // BF16 transposed result, or null
    int n_store;                 // columns stored (<= N)
    const unsigned long long* m_dev;  // when set, rows from *m_dev rounded up to 32 on are skipped
};
void launch_gemm(const GemmArgs& g, hipStream_t s);

// Layer 1 forward in FP32: z1 = x W1^T + b1, a1 = SiLU(z1) (x: [M][IN] FP32, W1 from
// the FP32 parameters); writes a1 BF16 [M][H] (and transposed [H][M] if a1t), z1 FP32.
// With m_dev set, rows from *m_dev rounded up to 32 on are skipped.
void launch_layer1_forward(const float* x, const float* P, int M, uint16_t* a1, uint16_t* a1t, float* z1, hipStream_t s,
                           const unsigned long long* m_dev = nullptr);
// dW1 = d1^T x over the M rows (FP32), db for every layer's delta (column sums).
void launch_layer1_wgrad(const float* d1, const float* x, int M, float* G, hipStream_t s);
void launch_colsum(const float* d, int M, int N, int ld, float* out, hipStream_t s);
// Adam step over the NPARAM FP32 parameters (t: step count from 1), then BF16 publish.
void launch_adam(float* P, const float* G, float* m1, float* m2, int t, float lr, hipStream_t s);
void launch_publish(const float* P, uint16_t* Wb, hipStream_t s);

// Loss of a training batch. out: [M][OUT] FP32 network outputs; T: [M][8] targets
// (y_harmonic, v, t, mask_v, mask_t, tau_t, 0, 0). Writes dO BF16 [M][OUT] and
// its transpose [OUT][M], an FP32 copy dOf [M][OUT], and adds the three loss sums into loss[3].
void launch_loss(const float* out, const float* T, int M, float w_t, uint16_t* dO, uint16_t* dOt, float* dOf, float* loss,
                 hipStream_t s);

}  // namespace ptg
