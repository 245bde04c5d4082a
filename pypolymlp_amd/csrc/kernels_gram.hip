// Gram accumulation C += X^T X (upper block triangle) on MFMA. The weighted
// targets sit in column F of X, so the same accumulation gives X^T y
// (C[:F, F]) and y^T y (C[F, F]) from exactly the X the Gram sees.
//
//   plain FP64 : v_mfma_f64_16x16x4_f64, FP64 accumulation (the reference's
//                numpy x.T @ x, src/pypolymlp/mlp_dev/core/utils_sequential.py:94-109).
//   mixed      : X in FP32, quantised per (row chunk of at most 4096 rows,
//                column) to a shared exponent E and three signed 7-bit
//                integer slices (the energy and stress rows in chunks of their
//                own, before the force rows),
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
// rows padded with zeros to a multiple of 64.
#include <hip/hip_runtime.h>

#include <stdexcept>

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
constexpr int kQChunk = 4096;  // rows sharing one exponent per column (at most)
constexpr int kQK = 32;        // rows per K step (one i8 MFMA)

// Row q of the quantised matrix is row rmap[q] of X (a zero row where
// rmap[q] < 0). The energy and stress rows of a sub-batch come first, then the
// force rows: an energy row sums a feature over all N atoms of its structure,
// so a chunk shared with force rows would set their exponent about N times too
// high. The Gram does not depend on the row order. Chunk 0 holds the first h0
// rows (a multiple of 64, at most kQChunk), chunk k > 0 the kQChunk rows from
// h0 + (k - 1) kQChunk, so the energy and stress rows end on a chunk boundary.
__host__ __device__ __forceinline__ int chunk_of(int q, int h0) {
    return q < h0 ? 0 : 1 + (q - h0) / kQChunk;
}
__host__ __device__ __forceinline__ int n_chunks_of(int rows, int h0) {
    return rows <= h0 ? 1 : 1 + (rows - h0 + kQChunk - 1) / kQChunk;
}
template <class T>
__device__ __forceinline__ T xrow(const T* __restrict__ X, const int* __restrict__ rmap, int q,
                                  int ld, int c) {
    const int r = rmap[q];
    return r >= 0 ? X[(size_t)r * ld + c] : T(0);
}
__device__ __forceinline__ int exponent_of(float m) {
    int e = 0;
    if (m > 0.f) (void)frexpf(m, &e);
    return e;
}
__device__ __forceinline__ int exponent_of(double m) {
    int e = 0;
    if (m > 0.0) (void)frexp(m, &e);
    return e;
}

// E[chunk][c]: smallest E with |x| < 2^E over the chunk's rows (0 for a zero column).
// 64 columns per workgroup, four row phases of 64 threads; each thread reads
// the row indices of its next eight rows, then keeps their eight loads in
// flight (the maximum is order independent, so the result does not depend on
// the split)
template <class T>
__global__ void __launch_bounds__(256) colexp_kernel(const T* __restrict__ X,
                                                     const int* __restrict__ rmap, int ld, int rows,
                                                     int h0, int ncols, int* __restrict__ E) {
    __shared__ T part[4][64];
    const int cx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int c = blockIdx.x * 64 + cx;
    const int chunk = blockIdx.y;
    const int r0 = chunk == 0 ? 0 : h0 + (chunk - 1) * kQChunk;
    const int r1 = min(rows, chunk == 0 ? h0 : r0 + kQChunk);
    T m[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) m[u] = T(0);
    if (c < ncols) {
        int r = r0 + ty;
        // the next batch's row indices load while this batch's rows of X do
        int rn[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) rn[u] = r + 28 < r1 ? rmap[r + 4 * u] : -1;
        for (; r + 28 < r1; r += 32) {
            int rr[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) rr[u] = rn[u];
            if (r + 60 < r1) {
#pragma unroll
                for (int u = 0; u < 8; ++u) rn[u] = rmap[r + 32 + 4 * u];
            }
            T v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = rr[u] >= 0 ? X[(size_t)rr[u] * ld + c] : T(0);
#pragma unroll
            for (int u = 0; u < 8; ++u) m[u] = fmax(m[u], fabs(v[u]));
        }
        for (; r < r1; r += 4) m[0] = fmax(m[0], fabs(xrow(X, rmap, r, ld, c)));
    }
#pragma unroll
    for (int u = 1; u < 8; ++u) m[0] = fmax(m[0], m[u]);
    part[ty][cx] = m[0];
    __syncthreads();
    if (ty == 0 && c < ncols) {
        const T m = fmax(fmax(part[0][cx], part[1][cx]), fmax(part[2][cx], part[3][cx]));
        E[(size_t)chunk * ld + c] = exponent_of(m);
    }
}

// Q[s][rb][c][k] (int8): slice s of x[rmap[32 rb + k]][c], k-contiguous per
// column so that 16 bytes are one lane's MFMA operand.
template <class T>
__global__ void quantize_kernel(const T* __restrict__ X, const int* __restrict__ rmap, int ld,
                                int rows_pad, int h0, int ncols, const int* __restrict__ E,
                                signed char* __restrict__ Q) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int rb = blockIdx.y;
    if (c >= ncols) return;
    const int nrb = rows_pad / kQK;
    const int e = E[(size_t)chunk_of(rb * kQK, h0) * ld + c];
    const size_t plane = (size_t)nrb * ld * kQK;
    signed char* q = Q + ((size_t)rb * ld + c) * kQK;
    // the block's 32 row indices (uniform) first, then all 32 loads of X in flight
    int rr[kQK];
#pragma unroll
    for (int k = 0; k < kQK; ++k) rr[k] = rmap[rb * kQK + k];
    T xv[kQK];
#pragma unroll
    for (int k = 0; k < kQK; ++k) xv[k] = rr[k] >= 0 ? X[(size_t)rr[k] * ld + c] : T(0);
#pragma unroll
    for (int k0 = 0; k0 < kQK; k0 += 16) {
        union {
            signed char b[16];
            v4i v;
        } d0, d1, d2;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const T x = xv[k0 + k];
            const T m = ldexp(x, 7 - e);  // |m| < 128
            const T t0 = trunc(m);
            const T r1 = (m - t0) * T(128);  // exact
            const T t1 = trunc(r1);
            const T r2 = (r1 - t1) * T(128);  // exact
            const T t2 = fmin(fmax(rint(r2), T(-127)), T(127));
            d0.b[k] = (signed char)t0, d1.b[k] = (signed char)t1, d2.b[k] = (signed char)t2;
        }
        *reinterpret_cast<v4i*>(q + k0) = d0.v;
        *reinterpret_cast<v4i*>(q + plane + k0) = d1.v;
        *reinterpret_cast<v4i*>(q + 2 * plane + k0) = d2.v;
    }
}

// ---------------------------------------------------------------------------
// Mixed: 128 (i) x 64 (j) output tile per 512-thread workgroup; each of the 8
// waves owns one 32x32 block with five int32 accumulators (levels s + t =
// 0..4), so two waves run per SIMD. Stages of 64 rows through two LDS buffers
// (48-byte column stride, b128 reads without bank conflicts). In each stage a
// wave reads its operands, stores the next stage (loaded into registers one
// stage earlier) into the other buffer and issues the loads of the stage after;
// the LDS and global instructions are interleaved with the 18 MFMAs. Per
// 4096-row chunk the int32 levels are combined into FP64 sums held in
// registers; C is read and written once, at the end.
// ---------------------------------------------------------------------------
constexpr int GI_TI = 128;
constexpr int GI_TJ = 64;
constexpr int GI_NT = 512;                       // threads per workgroup
constexpr int GI_CS = 48;                        // LDS bytes per column and 32-row block
constexpr int GI_RB = 2;                         // 32-row blocks per stage
constexpr int GI_A = 3 * GI_RB * GI_TI * GI_CS;  // A panel bytes per stage
constexpr int GI_B = 3 * GI_RB * GI_TJ * GI_CS;  // B panel bytes per stage
constexpr int GI_PA = 3 * GI_RB * GI_TI * 2;     // 16-byte A pieces per stage
constexpr int GI_PB = 3 * GI_RB * GI_TJ * 2;     // 16-byte B pieces per stage
constexpr int GI_LA = (GI_PA + GI_NT - 1) / GI_NT;  // A pieces per thread
constexpr int GI_LB = (GI_PB + GI_NT - 1) / GI_NT;  // B pieces per thread
constexpr int GI_MFMA = 9 * GI_RB;               // MFMAs per wave and stage
static_assert(GI_PA % 64 == 0 && GI_PB % 64 == 0 && GI_TJ % 32 == 0,
              "staging pieces come in runs of 64 (32 columns x 2 halves)");

// Tile order: workgroups are dealt to the 8 XCDs round robin; each XCD runs
// one super-tile of 4 x 8 tiles (512 x 512 outputs) at a time, so the panels
// of concurrently running tiles are shared in its L2. Super-tiles (SI, SJ),
// SJ >= SI, cover the upper triangle; tiles below it return at once.
__device__ __forceinline__ bool tile_index_i8(int b, int nbi, int& bi, int& bj) {
    const int xcd = b & 7, k = b >> 3;
    const int sup = (k >> 5) * 8 + xcd, loc = k & 31;
    const int nsi = (nbi + 3) / 4;
    int r = 0, rem = sup;
    while (r < nsi && rem >= nsi - r) rem -= nsi - r, ++r;
    if (r >= nsi) return false;
    bi = r * 4 + (loc >> 3);
    bj = (r + rem) * 8 + (loc & 7);
    return bi < nbi && bj < 2 * nbi && bj >= 2 * bi;
}

static long gram_i8_grid(int nbi) {
    const long nsi = (nbi + 3) / 4, nsup = nsi * (nsi + 1) / 2;
    return (nsup + 7) / 8 * 8 * 32;
}

size_t gram_i8_smem() { return 2 * (size_t)(GI_A + GI_B); }

__global__ void __launch_bounds__(GI_NT) gram_i8_kernel(const signed char* __restrict__ Q, int ld,
                                                        int rows_pad, int h0, int nbi,
                                                        const int* __restrict__ E,
                                                        double* __restrict__ C, int ldc) {
    extern __shared__ __attribute__((aligned(16))) signed char smem_i8[];
    int bi, bj;
    if (!tile_index_i8(blockIdx.x, nbi, bi, bj)) return;
    const int I0 = bi * GI_TI, J0 = bj * GI_TJ;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wi = (wave & 3) * 32, wj = (wave >> 2) * 32;
    const int nrb = rows_pad / kQK;
    const int nst = nrb / GI_RB;
    const size_t plane = (size_t)nrb * ld * kQK;
    v16i acc[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) acc[k] = v16i{};
    double sacc[16];
#pragma unroll
    for (int v = 0; v < 16; ++v) sacc[v] = 0.0;
    // piece p -> ((s * GI_RB + rbo) * cols + col) * 2 + half
    // within each run of 64 pieces, pieces 0-31 are the first 16-byte halves of
    // 32 consecutive columns and pieces 32-63 their second halves, so the 16
    // lanes of one LDS store pass write 16 columns of one half (distinct banks
    // at the 48-byte column stride, as the operand reads)
    auto piece = [&](int p, int cols, int& half, int& col, int& sr) {
        const int q = (p >> 6) * 32 + (p & 31);
        half = (p >> 5) & 1, col = q % cols, sr = q / cols;
    };
    auto g_off = [&](int p, int cols, int col0, int st) -> size_t {
        int half, col, sr;
        piece(p, cols, half, col, sr);
        const int s = sr / GI_RB, rbo = sr % GI_RB;
        return s * plane + ((size_t)(st * GI_RB + rbo) * ld + col0 + col) * kQK + 16 * half;
    };
    auto l_off = [&](int p, int cols) -> int {
        int half, col, sr;
        piece(p, cols, half, col, sr);
        return (sr * cols + col) * GI_CS + 16 * half;
    };
    // thread pieces past a panel's end wrap around (branch-free): such a
    // thread loads and stores a copy of a piece another thread stores, the
    // same value into the same slot
    v4i ra[GI_LA], rb[GI_LB];
    auto load = [&](int st) {
#pragma unroll
        for (int r = 0; r < GI_LA; ++r)
            ra[r] = *reinterpret_cast<const v4i*>(Q + g_off((tid + GI_NT * r) % GI_PA, GI_TI, I0, st));
#pragma unroll
        for (int r = 0; r < GI_LB; ++r)
            rb[r] = *reinterpret_cast<const v4i*>(Q + g_off((tid + GI_NT * r) % GI_PB, GI_TJ, J0, st));
    };
    auto store = [&](int buf) {
        signed char* A = smem_i8 + (size_t)buf * (GI_A + GI_B);
        signed char* B = A + GI_A;
#pragma unroll
        for (int r = 0; r < GI_LA; ++r)
            *reinterpret_cast<v4i*>(A + l_off((tid + GI_NT * r) % GI_PA, GI_TI)) = ra[r];
#pragma unroll
        for (int r = 0; r < GI_LB; ++r)
            *reinterpret_cast<v4i*>(B + l_off((tid + GI_NT * r) % GI_PB, GI_TJ)) = rb[r];
    };
    // x~_i x~_j = 2^(E_i + E_j - 14) sum_k acc_k 2^(-7k): exact int32 sums
    // combined in FP64; D[(v%4) + 8(v/4) + 4(l/32)][l%32]
    const int i0 = I0 + wi + 4 * (lane >> 5);
    const int j = J0 + wj + (lane & 31);
    auto flush = [&](int chunk) {
        const int* Ec = E + (size_t)chunk * ld;
        const int ej = Ec[j];
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int di = (v & 3) + 8 * (v >> 2);
            double t = double(acc[4][v]);
            t = t * 0.0078125 + double(acc[3][v]);
            t = t * 0.0078125 + double(acc[2][v]);
            t = t * 0.0078125 + double(acc[1][v]);
            t = t * 0.0078125 + double(acc[0][v]);
            sacc[v] += ldexp(t, Ec[i0 + di] + ej - 14);
        }
#pragma unroll
        for (int k = 0; k < 5; ++k) acc[k] = v16i{};
    };
    const int st_per_chunk = kQChunk / (kQK * GI_RB);
    const int h0s = h0 / (kQK * GI_RB);  // stages of chunk 0
    const int nchunk = nst <= h0s ? 1 : 1 + (nst - h0s + st_per_chunk - 1) / st_per_chunk;
    load(0);
    store(0);
    load(nst > 1 ? 1 : 0);
    __syncthreads();
    for (int ch = 0; ch < nchunk; ++ch) {
        const int s1 = min(nst, h0s + ch * st_per_chunk);
        for (int st = ch == 0 ? 0 : h0s + (ch - 1) * st_per_chunk; st < s1; ++st) {
            const signed char* A = smem_i8 + (size_t)(st & 1) * (GI_A + GI_B);
            const signed char* B = A + GI_A;
            // lane -> column (lane % 32) of its block, k half (lane / 32)
            v4i a[GI_RB][3], b[GI_RB][3];
#pragma unroll
            for (int rbo = 0; rbo < GI_RB; ++rbo)
#pragma unroll
                for (int s = 0; s < 3; ++s) {
                    a[rbo][s] = *reinterpret_cast<const v4i*>(
                        A + ((s * GI_RB + rbo) * GI_TI + wi + (lane & 31)) * GI_CS + 16 * (lane >> 5));
                    b[rbo][s] = *reinterpret_cast<const v4i*>(
                        B + ((s * GI_RB + rbo) * GI_TJ + wj + (lane & 31)) * GI_CS + 16 * (lane >> 5));
                }
            // stage st + 1 into the other buffer (last read in iteration st - 1),
            // then the loads of stage st + 2 (past the end: the last stage again)
            store((st + 1) & 1);
            load(min(st + 2, nst - 1));
            // issue order: the k-th product of every level before any (k+1)-th
            constexpr int kS[3][5] = {{0, 0, 0, 1, 2}, {-1, 1, 1, 2, -1}, {-1, -1, 2, -1, -1}};
#pragma unroll
            for (int rbo = 0; rbo < GI_RB; ++rbo)
#pragma unroll
                for (int k = 0; k < 3; ++k)
#pragma unroll
                    for (int lvl = 0; lvl < 5; ++lvl) {
                        const int s = kS[k][lvl];
                        if (s < 0) continue;
                        acc[lvl] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[rbo][s], b[rbo][lvl - s],
                                                                         acc[lvl], 0, 0, 0);
                    }
            // schedule: the first row block's operands, then one MFMA per LDS
            // read of the second row block, per LDS store, per global load
            __builtin_amdgcn_sched_group_barrier(0x100, 3 * 2, 0);
#pragma unroll
            for (int i = 0; i < 3 * 2; ++i) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
            }
#pragma unroll
            for (int i = 0; i < GI_LA + GI_LB; ++i) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);
            }
#pragma unroll
            for (int i = 0; i < GI_LA + GI_LB; ++i) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
            }
            __builtin_amdgcn_sched_group_barrier(0x008, GI_MFMA - 6 - 2 * (GI_LA + GI_LB), 0);
            __syncthreads();
        }
        flush(ch);
    }
    double* cb = C + (size_t)i0 * ldc + j;
#pragma unroll
    for (int v = 0; v < 16; ++v) cb[(size_t)((v & 3) + 8 * (v >> 2)) * ldc] += sacc[v];
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

size_t gram_quant_bytes(int rows_pad, int h0, int ld) {
    return 3 * (size_t)rows_pad * ld + (size_t)n_chunks_of(rows_pad, h0) * ld * sizeof(int);
}

void launch_gram_f64(hipStream_t st, const double* X, int ld, int rows_pad, int F_pad, double* C,
                     int ldc) {
    const int nb = F_pad / G64_T;
    gram_f64_kernel<<<nb * (nb + 1) / 2, 256, 0, st>>>(X, ld, rows_pad, nb, C, ldc);
}

int gram_chunk_rows() { return kQChunk; }

void launch_gram_quantize(hipStream_t st, bool x_fp64, const void* X, const int* rmap, int ld,
                          int rows_pad, int h0, int F_pad, void* quant) {
    if (h0 < kQK * GI_RB || h0 > kQChunk || h0 % (kQK * GI_RB) || rows_pad % (kQK * GI_RB))
        throw std::invalid_argument("quantisation chunk rows");
    const int n_chunks = n_chunks_of(rows_pad, h0);
    signed char* Q = reinterpret_cast<signed char*>(quant);
    int* E = reinterpret_cast<int*>(Q + 3 * (size_t)rows_pad * ld);
    const int B = 256;
    const dim3 ge((F_pad + 63) / 64, n_chunks), gq((F_pad + B - 1) / B, rows_pad / kQK);
    if (x_fp64) {
        const double* Xd = static_cast<const double*>(X);
        colexp_kernel<double><<<ge, 256, 0, st>>>(Xd, rmap, ld, rows_pad, h0, F_pad, E);
        quantize_kernel<double><<<gq, B, 0, st>>>(Xd, rmap, ld, rows_pad, h0, F_pad, E, Q);
    } else {
        const float* Xf = static_cast<const float*>(X);
        colexp_kernel<float><<<ge, 256, 0, st>>>(Xf, rmap, ld, rows_pad, h0, F_pad, E);
        quantize_kernel<float><<<gq, B, 0, st>>>(Xf, rmap, ld, rows_pad, h0, F_pad, E, Q);
    }
}

void launch_gram_i8(hipStream_t st, int ld, int rows_pad, int h0, int F_pad, const void* quant,
                    double* C, int ldc) {
    const signed char* Q = reinterpret_cast<const signed char*>(quant);
    const int* E = reinterpret_cast<const int*>(Q + 3 * (size_t)rows_pad * ld);
    const int nbi = F_pad / GI_TI;
    const size_t sm = gram_i8_smem();
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&gram_i8_kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    gram_i8_kernel<<<gram_i8_grid(nbi), GI_NT, sm, st>>>(Q, ld, rows_pad, h0, nbi, E, C, ldc);
}

}  // namespace plm
