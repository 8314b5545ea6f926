// Kernels of the surrogate's multilayer perceptron (layout and roles in kern_mlp.h).
#include "kern_mlp.h"

namespace ptg {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// FP32 -> BF16, round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}
__device__ __forceinline__ float sigmoid(float z) { return 1.0f / (1.0f + expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigmoid(z); }
__device__ __forceinline__ float silu_grad(float z) {
    float sg = sigmoid(z);
    return sg * (1.0f + z * (1.0f - sg));
}

// One wave per 32 x 32 tile of C. v_mfma_f32_32x32x8_bf16: lane l supplies
// A[l % 32][4 (l / 32) .. +3] and B[4 (l / 32) .. +3][l % 32] and receives
// D[8 (q / 4) + 4 (l / 32) + q % 4][l % 32] in accumulator q.
__global__ __launch_bounds__(64) void k_gemm_bf16(GemmArgs g) {
    const int lane = threadIdx.x;
    const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
    if (g.m_dev && uint64_t(i0) >= (*g.m_dev + 31) / 32 * 32) return;
    const int r = lane & 31, kq = (lane >> 5) * 4;
    const uint16_t* pa = g.A + size_t(i0 + r) * g.lda + kq;
    const uint16_t* pb = g.Bt + size_t(j0 + r) * g.ldb + kq;
    f32x16 acc;
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.0f;
    for (int k = 0; k < g.K; k += 8) {
        const bf16x4 a = *reinterpret_cast<const bf16x4*>(pa + k);
        const bf16x4 b = *reinterpret_cast<const bf16x4*>(pb + k);
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
    const int j = j0 + r;
    if (j >= g.n_store) return;
    const float bj = g.bias ? g.bias[j] : 0.0f;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int i = i0 + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
        float zv = acc[q] + bj, y = zv;
        if (g.act == 1) y = silu(zv);
        else if (g.act == 2) y = zv * silu_grad(g.zin[size_t(i) * g.ldzi + j]);
        if (g.z) g.z[size_t(i) * g.ldz + j] = g.act == 1 ? zv : y;
        if (g.C) g.C[size_t(i) * g.ldc + j] = to_bf16(y);
        if (g.Ct) g.Ct[size_t(j) * g.ldct + i] = to_bf16(y);
    }
}

void launch_gemm(const GemmArgs& g, hipStream_t s) {
    dim3 grid(g.M / 32, g.N / 32);
    if (g.M > 0 && g.N > 0) k_gemm_bf16<<<grid, 64, 0, s>>>(g);
}

// Layer 1 in FP32: one thread per (row, hidden unit).
__global__ void k_layer1_forward(const float* __restrict__ x, const float* __restrict__ P, int M, uint16_t* a1, uint16_t* a1t,
                                 float* z1, const unsigned long long* m_dev) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= M * mlp::H) return;
    const int i = idx / mlp::H, h = idx % mlp::H;
    if (m_dev && uint64_t(i) >= (*m_dev + 31) / 32 * 32) return;
    const float* w = P + mlp::W1 + h * mlp::IN;
    const float* xi = x + size_t(i) * mlp::IN;
    float zv = P[mlp::B1 + h];
#pragma unroll
    for (int k = 0; k < mlp::IN; ++k) zv = fmaf(xi[k], w[k], zv);
    const float y = silu(zv);
    if (z1) z1[idx] = zv;
    a1[idx] = to_bf16(y);
    if (a1t) a1t[size_t(h) * M + i] = to_bf16(y);
}
void launch_layer1_forward(const float* x, const float* P, int M, uint16_t* a1, uint16_t* a1t, float* z1, hipStream_t s,
                           const unsigned long long* m_dev) {
    const int n = M * mlp::H;
    if (n > 0) k_layer1_forward<<<(n + 255) / 256, 256, 0, s>>>(x, P, M, a1, a1t, z1, m_dev);
}

// dW1[h][k] = sum_i d1[i][h] x[i][k]: one block of 256 per (h, k), a tree sum over rows.
__global__ void k_layer1_wgrad(const float* __restrict__ d1, const float* __restrict__ x, int M, float* G) {
    const int h = blockIdx.x / mlp::IN, k = blockIdx.x % mlp::IN;
    __shared__ float red[256];
    float acc = 0.0f;
    for (int i = threadIdx.x; i < M; i += 256) acc = fmaf(d1[size_t(i) * mlp::H + h], x[size_t(i) * mlp::IN + k], acc);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) G[mlp::W1 + h * mlp::IN + k] = red[0];
}
void launch_layer1_wgrad(const float* d1, const float* x, int M, float* G, hipStream_t s) {
    k_layer1_wgrad<<<mlp::H * mlp::IN, 256, 0, s>>>(d1, x, M, G);
}

// out[n] = sum_i d[i][n] (bias gradients): one block of 256 per column.
__global__ void k_colsum(const float* __restrict__ d, int M, int ld, float* out) {
    const int n = blockIdx.x;
    __shared__ float red[256];
    float acc = 0.0f;
    for (int i = threadIdx.x; i < M; i += 256) acc += d[size_t(i) * ld + n];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[n] = red[0];
}
void launch_colsum(const float* d, int M, int N, int ld, float* out, hipStream_t s) {
    if (N > 0) k_colsum<<<N, 256, 0, s>>>(d, M, ld, out);
}

// Adam (beta1 0.9, beta2 0.999, eps 1e-8) with bias correction.
__global__ void k_adam(float* P, const float* __restrict__ G, float* m1, float* m2, int t, float lr) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= mlp::NPARAM) return;
    const float b1 = 0.9f, b2 = 0.999f;
    const float g = G[i];
    const float m = b1 * m1[i] + (1.0f - b1) * g;
    const float v = b2 * m2[i] + (1.0f - b2) * g * g;
    m1[i] = m;
    m2[i] = v;
    const float mh = m / (1.0f - powf(b1, float(t))), vh = v / (1.0f - powf(b2, float(t)));
    P[i] -= lr * mh / (sqrtf(vh) + 1e-8f);
}
void launch_adam(float* P, const float* G, float* m1, float* m2, int t, float lr, hipStream_t s) {
    k_adam<<<(mlp::NPARAM + 255) / 256, 256, 0, s>>>(P, G, m1, m2, t, lr);
}

// BF16 copies of W1..W4 and the transposes of W2..W4.
__global__ void k_publish(const float* __restrict__ P, uint16_t* Wb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    using namespace mlp;
    if (i < H * H) {
        const int r = i / H, c = i % H;
        Wb[BW2 + i] = to_bf16(P[W2 + i]);
        Wb[BW2T + c * H + r] = to_bf16(P[W2 + i]);
        Wb[BW3 + i] = to_bf16(P[W3 + i]);
        Wb[BW3T + c * H + r] = to_bf16(P[W3 + i]);
    }
    if (i < OUT * H) {
        const int r = i / H, c = i % H;
        Wb[BW4 + i] = to_bf16(P[W4 + i]);
        Wb[BW4T + c * OUT + r] = to_bf16(P[W4 + i]);
    }
    if (i < H * IN) Wb[BW1 + i] = to_bf16(P[W1 + i]);
}
void launch_publish(const float* P, uint16_t* Wb, hipStream_t s) {
    k_publish<<<(mlp::H * mlp::H + 255) / 256, 256, 0, s>>>(P, Wb);
}

// Loss: output 0 regresses the harmonic target (mean square), output 1 the
// rank-coupled lightfront fraction (mean square, rows with mask_v), output 2
// the angle by the pinball loss at quantile tau (rows with mask_t, weight w_t).
// Gradients are of the batch mean.
__global__ void k_loss(const float* __restrict__ out, const float* __restrict__ T, int M, float w_t, uint16_t* dO, uint16_t* dOt,
                       float* dOf, float* loss) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float l0 = 0.0f, l1 = 0.0f, l2 = 0.0f;
    if (i < M) {
        const float* o = out + size_t(i) * mlp::OUT;
        const float* t = T + size_t(i) * 8;
        const float inv = 1.0f / float(M);
        const float e0 = o[0] - t[0];
        const float e1 = (o[1] - t[1]) * t[3];
        const float y = t[2], tau = t[5], mt = t[4] * w_t;
        const float r2 = y - o[2];
        l0 = e0 * e0;
        l1 = e1 * e1;
        l2 = mt * (r2 >= 0.0f ? tau * r2 : (tau - 1.0f) * r2);
        const float g2 = mt * (r2 > 0.0f ? -tau : 1.0f - tau);
        for (int c = 0; c < mlp::OUT; ++c) {
            float g = c == 0 ? 2.0f * e0 * inv : (c == 1 ? 2.0f * e1 * inv : (c == 2 ? g2 * inv : 0.0f));
            dO[size_t(i) * mlp::OUT + c] = to_bf16(g);
            dOt[size_t(c) * M + i] = to_bf16(g);
            dOf[size_t(i) * mlp::OUT + c] = g;
        }
    }
    // wave sums, one atomic per wave and loss
    for (int off = 32; off > 0; off >>= 1) {
        l0 += __shfl_xor(l0, off, 64);
        l1 += __shfl_xor(l1, off, 64);
        l2 += __shfl_xor(l2, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(loss + 0, l0);
        atomicAdd(loss + 1, l1);
        atomicAdd(loss + 2, l2);
    }
}
void launch_loss(const float* out, const float* T, int M, float w_t, uint16_t* dO, uint16_t* dOt, float* dOf, float* loss,
                 hipStream_t s) {
    if (M > 0) k_loss<<<(M + 255) / 256, 256, 0, s>>>(out, T, M, w_t, dO, dOt, dOf, loss);
}

}  // namespace ptg
