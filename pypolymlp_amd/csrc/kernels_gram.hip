// Gram accumulation C += X^T X (upper block triangle) and X^T y on MFMA.
//
//   plain FP64 : v_mfma_f64_16x16x4_f64, FP64 accumulation (the reference's
//                numpy x.T @ x, src/mlp_dev/core/utils_sequential.py:94-109).
//   mixed      : X in FP32 split as x = hi + lo (two bf16), three products
//                hi*hi + hi*lo + lo*hi per element on v_mfma_f32_32x32x16_bf16
//                with FP32 accumulation over a row chunk, then added to the
//                FP64 result (block sums in FP64).
//
// X is row-major with leading dimension ld (>= F, multiple of the tile) and
// rows padded with zeros to a multiple of the K step.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace plm {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float floatx16;
typedef __attribute__((__vector_size__(4 * sizeof(double)))) double doublex4;

// Map a linear tile id to (bi, bj) with bi <= bj over nb blocks.
__device__ __forceinline__ void tri_index(int t, int nb, int& bi, int& bj) {
    // rows of the upper triangle: row bi has nb - bi tiles
    int r = 0, rem = t;
    while (rem >= nb - r) rem -= nb - r, ++r;
    bi = r, bj = r + rem;
}

// ---------------------------------------------------------------------------
// FP64: 64x64 output tile per 256-thread workgroup; each wave a 32x32 quarter
// as 2x2 MFMA 16x16 blocks; K step 16 rows through LDS.
// ---------------------------------------------------------------------------
constexpr int G64_T = 64;
constexpr int G64_K = 16;

__global__ void __launch_bounds__(256) gram_f64_kernel(const double* __restrict__ X, int ld,
                                                       int rows, int nb, double* __restrict__ C,
                                                       int ldc) {
    __shared__ double As[G64_K][G64_T + 1];
    __shared__ double Bs[G64_K][G64_T + 1];
    int bi, bj;
    tri_index(blockIdx.x, nb, bi, bj);
    const int I0 = bi * G64_T, J0 = bj * G64_T;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wi = (wave >> 1) * 32, wj = (wave & 1) * 32;
    doublex4 acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = doublex4{0, 0, 0, 0};
    for (int k0 = 0; k0 < rows; k0 += G64_K) {
        __syncthreads();
        for (int e = tid; e < G64_K * G64_T; e += 256) {
            const int kk = e / G64_T, c = e % G64_T;
            const size_t r = (size_t)(k0 + kk) * ld;
            As[kk][c] = X[r + I0 + c];
            Bs[kk][c] = X[r + J0 + c];
        }
        __syncthreads();
        for (int ks = 0; ks < G64_K; ks += 4) {
            const int kk = ks + (lane >> 4);
            double a[2], b[2];
            for (int m = 0; m < 2; ++m) a[m] = As[kk][wi + 16 * m + (lane & 15)];
            for (int n = 0; n < 2; ++n) b[n] = Bs[kk][wj + 16 * n + (lane & 15)];
            for (int m = 0; m < 2; ++m)
                for (int n = 0; n < 2; ++n)
                    acc[m][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[m], b[n], acc[m][n], 0, 0, 0);
        }
    }
    // D[l/16 + 4v][l%16] of each 16x16 block
    for (int m = 0; m < 2; ++m)
        for (int n = 0; n < 2; ++n)
            for (int v = 0; v < 4; ++v) {
                const int i = I0 + wi + 16 * m + (lane >> 4) + 4 * v;
                const int j = J0 + wj + 16 * n + (lane & 15);
                C[(size_t)i * ldc + j] += acc[m][n][v];
            }
}

// ---------------------------------------------------------------------------
// Mixed: 128x128 output tile per 256-thread workgroup (2x2 waves of 64x64 =
// 2x2 MFMA 32x32 blocks each), K step 32 rows; A/B panels staged in LDS as
// bf16 hi/lo, k contiguous per column.
// ---------------------------------------------------------------------------
constexpr int GB_T = 128;
constexpr int GB_K = 32;
constexpr int GB_KP = GB_K + 8;  // padded k stride (bf16 elements)

__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo) {
    hi = (__bf16)x;
    lo = (__bf16)(x - (float)hi);
}

__global__ void __launch_bounds__(256) gram_bf16x3_kernel(const float* __restrict__ X, int ld,
                                                          int rows, int nb, int chunk_rows,
                                                          double* __restrict__ C, int ldc) {
    __shared__ __attribute__((aligned(16))) __bf16 Ah[GB_T * GB_KP];
    __shared__ __attribute__((aligned(16))) __bf16 Al[GB_T * GB_KP];
    __shared__ __attribute__((aligned(16))) __bf16 Bh[GB_T * GB_KP];
    __shared__ __attribute__((aligned(16))) __bf16 Bl[GB_T * GB_KP];
    int bi, bj;
    tri_index(blockIdx.x, nb, bi, bj);
    const int I0 = bi * GB_T, J0 = bj * GB_T;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wi = (wave >> 1) * 64, wj = (wave & 1) * 64;
    floatx16 acc[2][2];
    auto zero = [&]() {
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.f;
    };
    auto flush = [&]() {
        for (int m = 0; m < 2; ++m)
            for (int n = 0; n < 2; ++n)
                for (int v = 0; v < 16; ++v) {
                    const int i = I0 + wi + 32 * m + (v & 3) + 8 * (v >> 2) + 4 * (lane >> 5);
                    const int j = J0 + wj + 32 * n + (lane & 31);
                    C[(size_t)i * ldc + j] += (double)acc[m][n][v];
                }
    };
    zero();
    int since = 0;
    for (int k0 = 0; k0 < rows; k0 += GB_K) {
        __syncthreads();
        // 32 rows x 128 cols per panel: each thread 16 consecutive columns of one row
        {
            const int kk = tid >> 3, cb = (tid & 7) * 16;
            const float* xr = X + (size_t)(k0 + kk) * ld;
            for (int c = 0; c < 16; c += 4) {
                const float4 va = *reinterpret_cast<const float4*>(xr + I0 + cb + c);
                const float4 vb = *reinterpret_cast<const float4*>(xr + J0 + cb + c);
                const float a4[4] = {va.x, va.y, va.z, va.w};
                const float b4[4] = {vb.x, vb.y, vb.z, vb.w};
                for (int u = 0; u < 4; ++u) {
                    __bf16 h, l;
                    split_bf16(a4[u], h, l);
                    Ah[(cb + c + u) * GB_KP + kk] = h;
                    Al[(cb + c + u) * GB_KP + kk] = l;
                    split_bf16(b4[u], h, l);
                    Bh[(cb + c + u) * GB_KP + kk] = h;
                    Bl[(cb + c + u) * GB_KP + kk] = l;
                }
            }
        }
        __syncthreads();
        for (int ks = 0; ks < GB_K; ks += 16) {
            const int kof = ks + 8 * (lane >> 5);
            bf16x8 ah[2], al[2], bh[2], bl[2];
            for (int m = 0; m < 2; ++m) {
                const int row = wi + 32 * m + (lane & 31);
                ah[m] = *reinterpret_cast<const bf16x8*>(&Ah[row * GB_KP + kof]);
                al[m] = *reinterpret_cast<const bf16x8*>(&Al[row * GB_KP + kof]);
            }
            for (int n = 0; n < 2; ++n) {
                const int col = wj + 32 * n + (lane & 31);
                bh[n] = *reinterpret_cast<const bf16x8*>(&Bh[col * GB_KP + kof]);
                bl[n] = *reinterpret_cast<const bf16x8*>(&Bl[col * GB_KP + kof]);
            }
            for (int m = 0; m < 2; ++m)
                for (int n = 0; n < 2; ++n) {
                    acc[m][n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[m], bh[n], acc[m][n], 0, 0, 0);
                    acc[m][n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[m], bl[n], acc[m][n], 0, 0, 0);
                    acc[m][n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[m], bh[n], acc[m][n], 0, 0, 0);
                }
        }
        since += GB_K;
        if (since >= chunk_rows) {
            flush();
            zero();
            since = 0;
        }
    }
    if (since > 0) flush();
}

// X^T y: xty[c] += sum_r X[r][c] * y[r]; one thread per column, rows split
// across blockIdx.y with FP64 atomics.
template <class T>
__global__ void xty_kernel(const T* __restrict__ X, int ld, int rows, int F,
                           const double* __restrict__ y, double* __restrict__ xty, int rows_per) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= F) return;
    const int r0 = blockIdx.y * rows_per, r1 = min(rows, r0 + rows_per);
    double s = 0.0;
    for (int r = r0; r < r1; ++r) s += double(X[(size_t)r * ld + c]) * y[r];
    atomicAdd(xty + c, s);
}

int gram_tile(bool fp64) { return fp64 ? G64_T : GB_T; }
int gram_kstep(bool fp64) { return fp64 ? G64_K : GB_K; }

void launch_gram(hipStream_t st, bool fp64, const void* X, int ld, int rows, int F_pad, double* C,
                 int ldc, int chunk_rows) {
    const int T = gram_tile(fp64);
    const int nb = F_pad / T;
    const int ntiles = nb * (nb + 1) / 2;
    if (fp64)
        gram_f64_kernel<<<ntiles, 256, 0, st>>>(reinterpret_cast<const double*>(X), ld, rows, nb, C,
                                                ldc);
    else
        gram_bf16x3_kernel<<<ntiles, 256, 0, st>>>(reinterpret_cast<const float*>(X), ld, rows, nb,
                                                   chunk_rows, C, ldc);
}

void launch_xty(hipStream_t st, bool fp64, const void* X, int ld, int rows, int F, const double* y,
                double* xty) {
    const int B = 256;
    const int rows_per = 2048;
    dim3 grid((F + B - 1) / B, (rows + rows_per - 1) / rows_per);
    if (fp64)
        xty_kernel<double><<<grid, B, 0, st>>>(reinterpret_cast<const double*>(X), ld, rows, F, y,
                                               xty, rows_per);
    else
        xty_kernel<float><<<grid, B, 0, st>>>(reinterpret_cast<const float*>(X), ld, rows, F, y,
                                              xty, rows_per);
}

}  // namespace plm
