// Gram accumulation C += X^T X (upper block triangle) on MFMA. The weighted
// targets sit in column F of X, so the same accumulation gives X^T y
// (C[:F, F]) and y^T y (C[F, F]) from exactly the X the Gram sees.
//
//   plain FP64 : v_mfma_f64_16x16x4_f64, FP64 accumulation (the reference's
//                numpy x.T @ x, src/pypolymlp/mlp_dev/core/utils_sequential.py:94-109).
//   mixed      : X in FP32, quantised per (4096-row chunk, column) to a shared
//                exponent E and three signed 7-bit integer slices,
//                x~ = 2^(E-7) (d0 + d1 2^-7 + d2 2^-14).
//                All nine slice products run on v_mfma_i32_32x32x32_i8 with
//                exact int32 accumulation, one accumulator per level s+t, and
//                are combined in FP64 once per chunk. The result is the exact
//                Gram of x~ (positive semidefinite, consistent with X^T y),
//                which the ill-conditioned ridge problem needs: a Gram that
//                drops products or rounds partial sums is indefinite at the
//                scale of the ridge parameter.
//
// X is row-major with leading dimension ld (>= F + 1, multiple of 128) and
// rows padded with zeros to a multiple of 32.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace plm {

typedef __attribute__((__vector_size__(4 * sizeof(double)))) double doublex4;
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

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
// Mixed: quantisation
// ---------------------------------------------------------------------------
constexpr int kQChunk = 4096;  // rows sharing one exponent per column
constexpr int kQK = 32;        // rows per K step (one i8 MFMA)

// E[chunk][c]: smallest E with |x| < 2^E over the chunk's rows (0 for a zero column).
__global__ void colexp_kernel(const float* __restrict__ X, int ld, int rows, int ncols,
                              int* __restrict__ E) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int chunk = blockIdx.y;
    if (c >= ncols) return;
    const int r0 = chunk * kQChunk, r1 = min(rows, r0 + kQChunk);
    float m = 0.f;
    for (int r = r0; r < r1; ++r) m = fmaxf(m, fabsf(X[(size_t)r * ld + c]));
    int e = 0;
    if (m > 0.f) (void)frexpf(m, &e);
    E[(size_t)chunk * ld + c] = e;
}

// Q[s][rb][c][k] (int8): slice s of x[32 rb + k][c], k-contiguous per column so
// that 16 bytes are one lane's MFMA operand.
__global__ void quantize_kernel(const float* __restrict__ X, int ld, int rows_pad, int ncols,
                                const int* __restrict__ E, signed char* __restrict__ Q) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int rb = blockIdx.y;
    if (c >= ncols) return;
    const int nrb = rows_pad / kQK;
    const int e = E[(size_t)((rb * kQK) / kQChunk) * ld + c];
    const size_t plane = (size_t)nrb * ld * kQK;
    signed char* q = Q + ((size_t)rb * ld + c) * kQK;
    for (int k0 = 0; k0 < kQK; k0 += 16) {
        union {
            signed char b[16];
            v4i v;
        } d0, d1, d2;
        for (int k = 0; k < 16; ++k) {
            const float x = X[(size_t)(rb * kQK + k0 + k) * ld + c];
            const float m = ldexpf(x, 7 - e);  // |m| < 128
            const float t0 = truncf(m);
            const float r1 = (m - t0) * 128.f;  // exact
            const float t1 = truncf(r1);
            const float r2 = (r1 - t1) * 128.f;  // exact
            const float t2 = fminf(fmaxf(rintf(r2), -127.f), 127.f);
            d0.b[k] = (signed char)t0, d1.b[k] = (signed char)t1, d2.b[k] = (signed char)t2;
        }
        *reinterpret_cast<v4i*>(q + k0) = d0.v;
        *reinterpret_cast<v4i*>(q + plane + k0) = d1.v;
        *reinterpret_cast<v4i*>(q + 2 * plane + k0) = d2.v;
    }
}

// ---------------------------------------------------------------------------
// Mixed: 128 (i) x 64 (j) output tile per 256-thread workgroup, each wave a
// 32x64 strip as two 32x32 blocks with five int32 accumulators each (levels
// s + t = 0..4). Operand panels (3 slices x columns x 32 rows) are staged in
// LDS, double-buffered, with a 48-byte column stride (b128 reads without bank
// conflicts). Tiles (bi, bj) with bj >= 2 bi cover the upper triangle.
// ---------------------------------------------------------------------------
constexpr int GI_TI = 128;
constexpr int GI_TJ = 64;
constexpr int GI_CS = 48;  // LDS bytes per column (32 + pad)

__device__ __forceinline__ void tile_index_i8(int t, int nbi, int& bi, int& bj) {
    const int nbj = 2 * nbi;
    int r = 0, rem = t;
    while (rem >= nbj - 2 * r) rem -= nbj - 2 * r, ++r;
    bi = r, bj = 2 * r + rem;
}

__global__ void __launch_bounds__(256) gram_i8_kernel(const signed char* __restrict__ Q, int ld,
                                                      int rows_pad, int nbi,
                                                      const int* __restrict__ E,
                                                      double* __restrict__ C, int ldc) {
    __shared__ __attribute__((aligned(16))) signed char smA[2][3 * GI_TI * GI_CS];
    __shared__ __attribute__((aligned(16))) signed char smB[2][3 * GI_TJ * GI_CS];
    int bi, bj;
    tile_index_i8(blockIdx.x, nbi, bi, bj);
    const int I0 = bi * GI_TI, J0 = bj * GI_TJ;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wi = wave * 32;
    const int nrb = rows_pad / kQK;
    const size_t plane = (size_t)nrb * ld * kQK;
    v16i acc[2][5];
#pragma unroll
    for (int n = 0; n < 2; ++n)
#pragma unroll
        for (int k = 0; k < 5; ++k) acc[n][k] = v16i{};
    // global -> registers for one K step (16 B pieces): A 3 x 128 columns x 2,
    // B 3 x 64 columns x 2 (threads < 128)
    v4i ld_a[3], ld_b[3];
    const int col = tid >> 1, half = tid & 1;
    const bool has_b = tid < 2 * GI_TJ;
    auto load = [&](int rb) {
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            const signed char* base = Q + s * plane + (size_t)rb * ld * kQK;
            ld_a[s] = *reinterpret_cast<const v4i*>(base + (size_t)(I0 + col) * kQK + 16 * half);
            if (has_b)
                ld_b[s] = *reinterpret_cast<const v4i*>(base + (size_t)(J0 + col) * kQK + 16 * half);
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            *reinterpret_cast<v4i*>(&smA[buf][(s * GI_TI + col) * GI_CS + 16 * half]) = ld_a[s];
            if (has_b)
                *reinterpret_cast<v4i*>(&smB[buf][(s * GI_TJ + col) * GI_CS + 16 * half]) = ld_b[s];
        }
    };
    load(0);
    store(0);
    __syncthreads();
    const int steps_per_chunk = kQChunk / kQK;
    for (int rb = 0; rb < nrb; ++rb) {
        const int buf = rb & 1;
        if (rb + 1 < nrb) load(rb + 1);
        // operands: lane -> column (lane % 32) of its block, k half (lane / 32)
        v4i a[3], b[2][3];
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            a[s] = *reinterpret_cast<const v4i*>(
                &smA[buf][(s * GI_TI + wi + (lane & 31)) * GI_CS + 16 * (lane >> 5)]);
#pragma unroll
            for (int n = 0; n < 2; ++n)
                b[n][s] = *reinterpret_cast<const v4i*>(
                    &smB[buf][(s * GI_TJ + 32 * n + (lane & 31)) * GI_CS + 16 * (lane >> 5)]);
        }
#pragma unroll
        for (int n = 0; n < 2; ++n)
#pragma unroll
            for (int s = 0; s < 3; ++s)
#pragma unroll
                for (int t = 0; t < 3; ++t)
                    acc[n][s + t] =
                        __builtin_amdgcn_mfma_i32_32x32x32_i8(a[s], b[n][t], acc[n][s + t], 0, 0, 0);
        if ((rb + 1) % steps_per_chunk == 0 || rb + 1 == nrb) {
            // x~_i x~_j = 2^(E_i + E_j - 14) sum_k acc_k 2^(-7k): exact int32
            // sums combined in FP64; D[(v%4) + 8(v/4) + 4(l/32)][l%32]
            const int* Ec = E + (size_t)(rb / steps_per_chunk) * ld;
#pragma unroll
            for (int n = 0; n < 2; ++n) {
                const int j = J0 + 32 * n + (lane & 31);
                const int ej = Ec[j];
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int i = I0 + wi + (v & 3) + 8 * (v >> 2) + 4 * (lane >> 5);
                    double t = double(acc[n][4][v]);
                    t = t * 0.0078125 + double(acc[n][3][v]);
                    t = t * 0.0078125 + double(acc[n][2][v]);
                    t = t * 0.0078125 + double(acc[n][1][v]);
                    t = t * 0.0078125 + double(acc[n][0][v]);
                    C[(size_t)i * ldc + j] += ldexp(t, Ec[i] + ej - 14);
                }
#pragma unroll
                for (int k = 0; k < 5; ++k) acc[n][k] = v16i{};
            }
        }
        if (rb + 1 < nrb) store(buf ^ 1);
        __syncthreads();
    }
}

// y (FP64, one value per row) into column col of X (FP32 or FP64)
template <class T>
__global__ void set_column_kernel(T* __restrict__ X, int ld, int col, const double* __restrict__ y,
                                  int rows) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < rows) X[(size_t)r * ld + col] = T(y[r]);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void launch_set_column(hipStream_t st, bool fp64, void* X, int ld, int col, const double* y,
                       int rows) {
    const int B = 256;
    if (fp64)
        set_column_kernel<double><<<(rows + B - 1) / B, B, 0, st>>>(reinterpret_cast<double*>(X),
                                                                    ld, col, y, rows);
    else
        set_column_kernel<float><<<(rows + B - 1) / B, B, 0, st>>>(reinterpret_cast<float*>(X), ld,
                                                                   col, y, rows);
}

size_t gram_quant_bytes(int rows_pad, int ld) {
    const size_t n_chunks = ((size_t)rows_pad + kQChunk - 1) / kQChunk;
    return 3 * (size_t)rows_pad * ld + n_chunks * ld * sizeof(int);
}

void launch_gram_f64(hipStream_t st, const double* X, int ld, int rows_pad, int F_pad, double* C,
                     int ldc) {
    const int nb = F_pad / G64_T;
    gram_f64_kernel<<<nb * (nb + 1) / 2, 256, 0, st>>>(X, ld, rows_pad, nb, C, ldc);
}

void launch_gram_quantize(hipStream_t st, const float* X, int ld, int rows_pad, int F_pad,
                          void* quant) {
    const int n_chunks = (rows_pad + kQChunk - 1) / kQChunk;
    signed char* Q = reinterpret_cast<signed char*>(quant);
    int* E = reinterpret_cast<int*>(Q + 3 * (size_t)rows_pad * ld);
    const int B = 256;
    colexp_kernel<<<dim3((F_pad + B - 1) / B, n_chunks), B, 0, st>>>(X, ld, rows_pad, F_pad, E);
    quantize_kernel<<<dim3((F_pad + B - 1) / B, rows_pad / kQK), B, 0, st>>>(X, ld, rows_pad,
                                                                              F_pad, E, Q);
}

void launch_gram_i8(hipStream_t st, int ld, int rows_pad, int F_pad, const void* quant, double* C,
                    int ldc) {
    const signed char* Q = reinterpret_cast<const signed char*>(quant);
    const int* E = reinterpret_cast<const int*>(Q + 3 * (size_t)rows_pad * ld);
    const int nbi = F_pad / GI_TI;
    gram_i8_kernel<<<nbi * (nbi + 1), 256, 0, st>>>(Q, ld, rows_pad, nbi, E, C, ldc);
}

}  // namespace plm
