// Surrogate network of the hyperparameter search: an FP32 multilayer
// perceptron 5-64-64-2 with SiLU activations, trained with Adam (decoupled
// weight decay) on a full batch. One workgroup runs every training step in a
// single launch with the parameters, Adam moments, activations and gradients
// resident in LDS.
//
// Parameter vector (row-major weights): [W1 (64x5) | b1 | W2 (64x64) | b2 |
// W3 (2x64) | b3]. Loss: mean over samples and outputs of (y - t)^2.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace plm {

constexpr int SI = kSurrIn, SH = kSurrHidden, SO = kSurrOut;
constexpr int OW1 = 0, OB1 = OW1 + SH * SI, OW2 = OB1 + SH, OB2 = OW2 + SH * SH,
              OW3 = OB2 + SH, OB3 = OW3 + SO * SH, NPAR = OB3 + SO;
static_assert(NPAR == kSurrParams, "parameter layout");

__device__ __forceinline__ float sigm(float z) { return 1.f / (1.f + expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float silu_d(float z) {
    const float s = sigm(z);
    return s * (1.f + z * (1.f - s));
}

// LDS: P, G (NPAR each); X (n x SI), T (n x SO), Z1, Z2, D (n x SH), Y (n x SO)
struct SurrLds {
    float *P, *G, *X, *T, *Z1, *Z2, *D, *Y;
    __device__ SurrLds(float* base, int n) {
        P = base;
        G = P + NPAR;
        X = G + NPAR;
        T = X + n * SI;
        Z1 = T + n * SO;
        Z2 = Z1 + n * SH;
        D = Z2 + n * SH;
        Y = D + n * SH;
    }
};

__device__ void surr_forward(const SurrLds& L, int n) {
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int e = tid; e < n * SH; e += nt) {
        const int s = e / SH, h = e % SH;
        float z = L.P[OB1 + h];
        for (int i = 0; i < SI; ++i) z += L.P[OW1 + h * SI + i] * L.X[s * SI + i];
        L.Z1[e] = z;
    }
    __syncthreads();
    for (int e = tid; e < n * SH; e += nt) {
        const int s = e / SH, h = e % SH;
        float z = L.P[OB2 + h];
        for (int k = 0; k < SH; ++k) z += L.P[OW2 + h * SH + k] * silu(L.Z1[s * SH + k]);
        L.Z2[e] = z;
    }
    __syncthreads();
    for (int e = tid; e < n * SO; e += nt) {
        const int s = e / SO, o = e % SO;
        float y = L.P[OB3 + o];
        for (int k = 0; k < SH; ++k) y += L.P[OW3 + o * SH + k] * silu(L.Z2[s * SH + k]);
        L.Y[e] = y;
    }
    __syncthreads();
}

// gradients of the loss into L.G; returns the loss (valid in every thread)
__device__ float surr_backward(const SurrLds& L, int n, float* red) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const float sc = 2.f / float(n * SO);
    // loss
    float part = 0.f;
    for (int e = tid; e < n * SO; e += nt) {
        const float d = L.Y[e] - L.T[e];
        part += d * d;
    }
    red[tid] = part;
    __syncthreads();
    if (tid == 0) {
        float s = 0.f;
        for (int k = 0; k < nt; ++k) s += red[k];
        red[nt] = s / float(n * SO);
    }
    // dY stored in place of Y (after the loss is read)
    __syncthreads();
    const float loss = red[nt];
    for (int e = tid; e < n * SO; e += nt) L.Y[e] = sc * (L.Y[e] - L.T[e]);
    __syncthreads();
    // W3, b3
    for (int e = tid; e < SO * SH; e += nt) {
        const int o = e / SH, k = e % SH;
        float g = 0.f;
        for (int s = 0; s < n; ++s) g += L.Y[s * SO + o] * silu(L.Z2[s * SH + k]);
        L.G[OW3 + e] = g;
    }
    for (int o = tid; o < SO; o += nt) {
        float g = 0.f;
        for (int s = 0; s < n; ++s) g += L.Y[s * SO + o];
        L.G[OB3 + o] = g;
    }
    // dZ2 = (dY W3) * silu'(Z2)
    for (int e = tid; e < n * SH; e += nt) {
        const int s = e / SH, k = e % SH;
        float d = 0.f;
        for (int o = 0; o < SO; ++o) d += L.Y[s * SO + o] * L.P[OW3 + o * SH + k];
        L.D[e] = d * silu_d(L.Z2[e]);
    }
    __syncthreads();
    // W2, b2
    for (int e = tid; e < SH * SH; e += nt) {
        const int h = e / SH, k = e % SH;
        float g = 0.f;
        for (int s = 0; s < n; ++s) g += L.D[s * SH + h] * silu(L.Z1[s * SH + k]);
        L.G[OW2 + e] = g;
    }
    for (int h = tid; h < SH; h += nt) {
        float g = 0.f;
        for (int s = 0; s < n; ++s) g += L.D[s * SH + h];
        L.G[OB2 + h] = g;
    }
    // dZ1 = (dZ2 W2) * silu'(Z1), written over Z2 (no longer needed)
    for (int e = tid; e < n * SH; e += nt) {
        const int s = e / SH, k = e % SH;
        float d = 0.f;
        for (int h = 0; h < SH; ++h) d += L.D[s * SH + h] * L.P[OW2 + h * SH + k];
        L.Z2[e] = d * silu_d(L.Z1[e]);
    }
    __syncthreads();
    // W1, b1
    for (int e = tid; e < SH * SI; e += nt) {
        const int h = e / SI, i = e % SI;
        float g = 0.f;
        for (int s = 0; s < n; ++s) g += L.Z2[s * SH + h] * L.X[s * SI + i];
        L.G[OW1 + e] = g;
    }
    for (int h = tid; h < SH; h += nt) {
        float g = 0.f;
        for (int s = 0; s < n; ++s) g += L.Z2[s * SH + h];
        L.G[OB1 + h] = g;
    }
    __syncthreads();
    return loss;
}

// mode 0: `steps` Adam steps (t counts from t0 + 1), loss of the last step;
// mode 1: gradient and loss at P (no update).
__global__ void __launch_bounds__(256) surr_train_kernel(float* __restrict__ P, float* __restrict__ M,
                                                         float* __restrict__ V,
                                                         const float* __restrict__ X,
                                                         const float* __restrict__ T, int n,
                                                         int steps, int t0, float lr, float beta1,
                                                         float beta2, float eps, float wd, int mode,
                                                         float* __restrict__ grad,
                                                         float* __restrict__ loss_out) {
    extern __shared__ float sm[];
    __shared__ float red[257];
    SurrLds L(sm, n);
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int e = tid; e < NPAR; e += nt) L.P[e] = P[e];
    for (int e = tid; e < n * SI; e += nt) L.X[e] = X[e];
    for (int e = tid; e < n * SO; e += nt) L.T[e] = T[e];
    __syncthreads();
    float loss = 0.f;
    if (mode == 1) {
        surr_forward(L, n);
        loss = surr_backward(L, n, red);
        for (int e = tid; e < NPAR; e += nt) grad[e] = L.G[e];
        if (tid == 0) *loss_out = loss;
        return;
    }
    // Adam moments stay in registers: each thread owns parameters tid + k*nt
    constexpr int kPer = (NPAR + 255) / 256;
    float m[kPer], v[kPer];
    for (int k = 0; k < kPer; ++k) {
        const int e = tid + k * nt;
        m[k] = e < NPAR ? M[e] : 0.f;
        v[k] = e < NPAR ? V[e] : 0.f;
    }
    for (int it = 0; it < steps; ++it) {
        surr_forward(L, n);
        loss = surr_backward(L, n, red);
        const int t = t0 + it + 1;
        const float c1 = 1.f - powf(beta1, float(t)), c2 = 1.f - powf(beta2, float(t));
        for (int k = 0; k < kPer; ++k) {
            const int e = tid + k * nt;
            if (e >= NPAR) continue;
            const float g = L.G[e];
            m[k] = beta1 * m[k] + (1.f - beta1) * g;
            v[k] = beta2 * v[k] + (1.f - beta2) * g * g;
            // decoupled weight decay (AdamW)
            L.P[e] -= lr * ((m[k] / c1) / (sqrtf(v[k] / c2) + eps) + wd * L.P[e]);
        }
        __syncthreads();
    }
    for (int k = 0; k < kPer; ++k) {
        const int e = tid + k * nt;
        if (e < NPAR) P[e] = L.P[e], M[e] = m[k], V[e] = v[k];
    }
    if (tid == 0) *loss_out = loss;
}

__global__ void __launch_bounds__(256) surr_predict_kernel(const float* __restrict__ P,
                                                           const float* __restrict__ X, int n,
                                                           float* __restrict__ Y) {
    extern __shared__ float sm[];
    SurrLds L(sm, n);
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int e = tid; e < NPAR; e += nt) L.P[e] = P[e];
    for (int e = tid; e < n * SI; e += nt) L.X[e] = X[e];
    __syncthreads();
    surr_forward(L, n);
    for (int e = tid; e < n * SO; e += nt) Y[e] = L.Y[e];
}

size_t surrogate_smem(int n) {
    return sizeof(float) * (2 * (size_t)NPAR + (size_t)n * (SI + 2 * SO + 3 * SH));
}

void launch_surr_train(hipStream_t st, float* P, float* M, float* V, const float* X,
                       const float* T, int n, int steps, int t0, float lr, float beta1,
                       float beta2, float eps, float wd, int mode, float* grad, float* loss) {
    const size_t sm = surrogate_smem(n);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&surr_train_kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    surr_train_kernel<<<1, 256, sm, st>>>(P, M, V, X, T, n, steps, t0, lr, beta1, beta2, eps, wd,
                                          mode, grad, loss);
}

void launch_surr_predict(hipStream_t st, const float* P, const float* X, int n, float* Y) {
    const size_t sm = surrogate_smem(n);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&surr_predict_kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    surr_predict_kernel<<<1, 256, sm, st>>>(P, X, n, Y);
}

}  // namespace plm
