// FP64 ridge solve on gfx950: blocked right-looking Cholesky A = L L^T
// (diagonal blocks in LDS, panel TRSM, trailing SYRK on FP64 MFMA),
// blocked triangular solves, and the residual GEMV used for one step of
// iterative refinement. Replaces LAPACK posv in solver_ridge
// (src/pypolymlp/mlp_dev/standard/solvers.py:9-84).
//
// Matrices are row-major with leading dimension ld; only the lower triangle
// is read and written by the factorization.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace plm {

constexpr int SB = 64;  // block size

typedef __attribute__((__vector_size__(4 * sizeof(double)))) double doublex4;

// Unblocked Cholesky of the diagonal block at (k0, k0) by one wave, in LDS
// (coalesced load and store). Column j, left-looking: lane i >= j forms
// a_ij - sum_{k<j} l_ik l_jk (k ascending: the same operations in the same
// order as the right-looking rank-1 updates), then divides by the pivot.
__global__ void __launch_bounds__(SB) potf2_kernel(double* __restrict__ A, int ld, int k0,
                                                   int* __restrict__ info) {
    __shared__ double a[SB][SB + 1];
    const int i = threadIdx.x;
    for (int e = i; e < SB * SB; e += SB) {
        const int r = e / SB, c = e % SB;
        a[r][c] = (c <= r) ? A[(size_t)(k0 + r) * ld + k0 + c] : 0.0;
    }
    __syncthreads();
    for (int j = 0; j < SB; ++j) {
        double v = 0.0;
        if (i >= j) {
            v = a[i][j];
#pragma unroll 8
            for (int k = 0; k < j; ++k) v -= a[i][k] * a[j][k];
            if (i == j) a[j][j] = v;
        }
        __syncthreads();
        double d = a[j][j];
        if (!(d > 0.0)) {
            if (i == j && *info == 0) *info = k0 + j + 1;
            d = 1.0;  // keep the factor finite; the caller discards this solve
        }
        const double djj = sqrt(d);
        __syncthreads();
        if (i == j) a[j][j] = djj;
        if (i > j) a[i][j] = v / djj;
        __syncthreads();
    }
    for (int e = i; e < SB * SB; e += SB) {
        const int r = e / SB, c = e % SB;
        if (c <= r) A[(size_t)(k0 + r) * ld + k0 + c] = a[r][c];
    }
}

// Panel: rows i in [k0+SB, n): A[i, k0:k0+SB] <- A[i, k0:k0+SB] L_kk^{-T} by
// forward substitution, one thread per row, the diagonal block in LDS; the
// loops are unrolled so that the row stays in registers.
__global__ void __launch_bounds__(256) trsm_panel_kernel(double* __restrict__ A, int ld, int k0,
                                                         int n) {
    __shared__ double L[SB][SB + 1];
    for (int e = threadIdx.x; e < SB * SB; e += 256) {
        const int i = e / SB, j = e % SB;
        L[i][j] = (j <= i) ? A[(size_t)(k0 + i) * ld + k0 + j] : 0.0;
    }
    __syncthreads();
    const int i = k0 + SB + blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double x[SB];
    double* row = A + (size_t)i * ld + k0;
#pragma unroll
    for (int c = 0; c < SB; ++c) x[c] = row[c];
#pragma unroll
    for (int c = 0; c < SB; ++c) {
        double v = x[c];
#pragma unroll
        for (int cp = 0; cp < c; ++cp) v -= x[cp] * L[c][cp];
        x[c] = v / L[c][c];
    }
#pragma unroll
    for (int c = 0; c < SB; ++c) row[c] = x[c];
}

// Update with a kw-wide panel (kw a multiple of SB): for the lower tiles
// (bi >= bj, bj < ncb) of the nt x nt block grid starting at (base, base),
// C -= P_bi P_bj^T with P = A[:, k0:k0+kw], on FP64 MFMA; the K loop runs over
// the panel in 64-column chunks through LDS and C is read and written once.
// ncb = nt: the whole trailing lower triangle; ncb < nt: only the first ncb
// column blocks (the rest of the current panel during its factorization).
__global__ void __launch_bounds__(256) syrk_panel_kernel(double* __restrict__ A, int ld, int k0,
                                                         int kw, int base, int nt) {
    __shared__ double Ps[SB][SB + 1];  // [k][i] panel rows of the I tile
    __shared__ double Qs[SB][SB + 1];  // [k][j] panel rows of the J tile
    int bj = 0, rem = blockIdx.x;
    while (rem >= nt - bj) rem -= nt - bj, ++bj;
    const int bi = bj + rem;
    const int I0 = base + bi * SB, J0 = base + bj * SB;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wi = (wave >> 1) * 32, wj = (wave & 1) * 32;
    doublex4 acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = doublex4{0, 0, 0, 0};
    for (int kc = 0; kc < kw; kc += SB) {
        __syncthreads();
        for (int e = tid; e < SB * SB; e += 256) {
            const int r = e / SB, k = e % SB;
            Ps[k][r] = A[(size_t)(I0 + r) * ld + k0 + kc + k];
            Qs[k][r] = A[(size_t)(J0 + r) * ld + k0 + kc + k];
        }
        __syncthreads();
        for (int ks = 0; ks < SB; ks += 4) {
            const int kk = ks + (lane >> 4);
            double av[2], bv[2];
            for (int m = 0; m < 2; ++m) av[m] = Ps[kk][wi + 16 * m + (lane & 15)];
            for (int q = 0; q < 2; ++q) bv[q] = Qs[kk][wj + 16 * q + (lane & 15)];
            for (int m = 0; m < 2; ++m)
                for (int q = 0; q < 2; ++q)
                    acc[m][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[m], bv[q], acc[m][q], 0, 0, 0);
        }
    }
    for (int m = 0; m < 2; ++m)
        for (int q = 0; q < 2; ++q)
            for (int v = 0; v < 4; ++v) {
                const int i = I0 + wi + 16 * m + (lane >> 4) + 4 * v;
                const int j = J0 + wj + 16 * q + (lane & 15);
                if (j <= i) A[(size_t)i * ld + j] -= acc[m][q][v];
            }
}

// Forward substitution on one diagonal block, z_k = L_kk^{-1} z_k, one wave:
// step c finishes z_c, then every lane below subtracts its term (each z_i sees
// the same operations in the same order as row-wise substitution).
__global__ void __launch_bounds__(SB) trsv_lower_block(const double* __restrict__ L, int ld,
                                                       int k0, double* __restrict__ z) {
    __shared__ double Lb[SB][SB + 1];
    __shared__ double zb[SB];
    const int t = threadIdx.x;
    for (int j = 0; j <= t; ++j) Lb[t][j] = L[(size_t)(k0 + t) * ld + k0 + j];
    zb[t] = z[k0 + t];
    __syncthreads();
    for (int c = 0; c < SB; ++c) {
        if (t == c) zb[c] /= Lb[c][c];
        __syncthreads();
        if (t > c) zb[t] -= Lb[t][c] * zb[c];
        __syncthreads();
    }
    z[k0 + t] = zb[t];
}

// Backward substitution on one diagonal block, x_k = L_kk^{-T} x_k, one wave.
__global__ void __launch_bounds__(SB) trsv_upper_block(const double* __restrict__ L, int ld,
                                                       int k0, double* __restrict__ z) {
    __shared__ double Lb[SB][SB + 1];
    __shared__ double zb[SB];
    const int t = threadIdx.x;
    for (int j = 0; j <= t; ++j) Lb[t][j] = L[(size_t)(k0 + t) * ld + k0 + j];
    zb[t] = z[k0 + t];
    __syncthreads();
    for (int c = SB - 1; c >= 0; --c) {
        if (t == c) zb[c] /= Lb[c][c];
        __syncthreads();
        if (t < c) zb[t] -= Lb[c][t] * zb[c];
        __syncthreads();
    }
    z[k0 + t] = zb[t];
}

// z[i] -= sum_c L[i][k0+c] z[k0+c] for rows i >= k0+SB (after a forward block).
__global__ void gemv_lower_update(const double* __restrict__ L, int ld, int k0, int n,
                                  double* __restrict__ z) {
    const int i = k0 + SB + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* row = L + (size_t)i * ld + k0;
    double s = 0.0;
    for (int c = 0; c < SB; ++c) s += row[c] * z[k0 + c];
    z[i] -= s;
}

// z[c] -= sum_r L[k0+r][c] x[k0+r] for columns c < k0 (after a backward block).
__global__ void gemv_upper_update(const double* __restrict__ L, int ld, int k0,
                                  double* __restrict__ z) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= k0) return;
    double s = 0.0;
    for (int r = 0; r < SB; ++r) s += L[(size_t)(k0 + r) * ld + c] * z[k0 + r];
    z[c] -= s;
}

// r = b - (A + alpha I) x with the full symmetric A: one wave per row, lanes
// over the row (coalesced), FP64 with compensated (TwoSum) accumulation per
// lane and a compensated combination of the lane sums.
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

__global__ void __launch_bounds__(256) residual_kernel(const double* __restrict__ A, int ld, int n,
                                                       double alpha, const double* __restrict__ x,
                                                       const double* __restrict__ b,
                                                       double* __restrict__ r) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= n) return;
    const double* row = A + (size_t)i * ld;
    double s = 0.0, comp = 0.0;
    for (int c = lane; c < n; c += 64) {
        double t, e;
        two_sum(s, row[c] * x[c], t, e);
        s = t, comp += e;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const double so = __shfl_xor(s, o), co = __shfl_xor(comp, o);
        double t, e;
        two_sum(s, so, t, e);
        s = t, comp += co + e;
    }
    if (lane == 0) r[i] = b[i] - ((s + comp) + alpha * x[i]);
}

__global__ void add_diag_kernel(double* __restrict__ A, int ld, int n, double add) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) A[(size_t)i * ld + i] += add;
}

__global__ void axpy_kernel(int n, const double* __restrict__ d, double* __restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += d[i];
}

// The scaled normal equations straight from the accumulator's Gram of
// [X | y] (upper triangle valid, leading dimension ldg; column F holds X^T y):
// A[i][j] = G[i][j] / s[i] / s[j] (the reference's xtx /= s[:, None];
// xtx /= s[None, :]), zero for dropped columns, identity on the padding
// (i, j >= F), the full symmetric n_pad x n_pad matrix; b[i] = G[i][F] / s[i].
// 32 x 32 tiles; the lower tiles read the mirrored upper tile through LDS.
__global__ void __launch_bounds__(256) gram_system_kernel(const double* __restrict__ G, int ldg,
                                                          int F, const double* __restrict__ s,
                                                          const unsigned char* __restrict__ zero,
                                                          double* __restrict__ A, int lda,
                                                          double* __restrict__ b) {
    __shared__ double T[32][33];
    const int bi = blockIdx.y, bj = blockIdx.x;
    const int r0 = min(bi, bj) * 32, c0 = max(bi, bj) * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int k = ty; k < 32; k += 8) {
        const int r = r0 + k, c = c0 + tx;
        T[k][tx] = (r < F && c < F) ? G[(size_t)r * ldg + c] : 0.0;
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int i = bi * 32 + k, j = bj * 32 + tx;
        double v;
        if (i >= F || j >= F) {
            v = i == j ? 1.0 : 0.0;
        } else if (zero[i] || zero[j]) {
            v = 0.0;
        } else {
            // element (min(i, j), max(i, j)) of the upper triangle
            const bool upper = bi < bj || (bi == bj && k <= tx);
            v = (upper ? T[k][tx] : T[tx][k]) / s[i] / s[j];
        }
        A[(size_t)i * lda + j] = v;
    }
    if (bj == 0 && threadIdx.x < 32) {
        const int i = bi * 32 + threadIdx.x;
        b[i] = (i < F && !zero[i]) ? G[(size_t)i * ldg + F] / s[i] : 0.0;
    }
}

// y = A x for the full symmetric n x n system (one wave per row, lanes over
// the row: coalesced), FP64
__global__ void __launch_bounds__(256) symv_rows_kernel(const double* __restrict__ A, int lda,
                                                        int n, const double* __restrict__ x,
                                                        double* __restrict__ y) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= n) return;
    const double* a = A + (size_t)row * lda;
    double acc = 0.0;
    for (int c = lane; c < n; c += 64) acc += a[c] * x[c];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (lane == 0) y[row] = acc;
}

void launch_gram_system(hipStream_t st, const double* G, int ldg, int F, const double* s,
                        const unsigned char* zero, double* A, int lda, int n_pad, double* b) {
    const dim3 grid(n_pad / 32, n_pad / 32);
    gram_system_kernel<<<grid, 256, 0, st>>>(G, ldg, F, s, zero, A, lda, b);
}

void launch_symv(hipStream_t st, const double* A, int lda, int n, const double* x, double* y) {
    symv_rows_kernel<<<(n + 3) / 4, 256, 0, st>>>(A, lda, n, x, y);
}

// Blocked right-looking Cholesky with NB-wide outer panels: inside a panel,
// 64-wide steps (diagonal block, panel TRSM below it, update of the rest of
// the panel's columns); after the panel, one trailing update with K = NB, so
// that the dominant product runs with an 8x wider inner dimension and the
// trailing matrix is read and written once per 512 columns.
constexpr int NB = 512;

void launch_cholesky(hipStream_t st, double* A, int ld, int n_pad, int* info) {
    for (int k0 = 0; k0 < n_pad; k0 += NB) {
        const int kw = min(NB, n_pad - k0);
        for (int j0 = k0; j0 < k0 + kw; j0 += SB) {
            potf2_kernel<<<1, SB, 0, st>>>(A, ld, j0, info);
            const int rest = n_pad - j0 - SB;
            if (rest <= 0) break;
            trsm_panel_kernel<<<(rest + 255) / 256, 256, 0, st>>>(A, ld, j0, n_pad);
            const int pc = (k0 + kw - j0 - SB) / SB;  // panel column blocks still to update
            if (pc > 0) {
                const int nt = rest / SB;
                syrk_panel_kernel<<<pc * nt - pc * (pc - 1) / 2, 256, 0, st>>>(A, ld, j0, SB,
                                                                               j0 + SB, nt);
            }
        }
        const int rest = n_pad - k0 - kw;
        if (rest > 0) {
            const int nt = rest / SB;
            syrk_panel_kernel<<<nt * (nt + 1) / 2, 256, 0, st>>>(A, ld, k0, kw, k0 + kw, nt);
        }
    }
}

void launch_chol_solve(hipStream_t st, const double* L, int ld, int n_pad, double* z) {
    for (int k0 = 0; k0 < n_pad; k0 += SB) {
        trsv_lower_block<<<1, 64, 0, st>>>(L, ld, k0, z);
        const int rest = n_pad - k0 - SB;
        if (rest > 0) gemv_lower_update<<<(rest + 255) / 256, 256, 0, st>>>(L, ld, k0, n_pad, z);
    }
    for (int k0 = n_pad - SB; k0 >= 0; k0 -= SB) {
        trsv_upper_block<<<1, 64, 0, st>>>(L, ld, k0, z);
        if (k0 > 0) gemv_upper_update<<<(k0 + 255) / 256, 256, 0, st>>>(L, ld, k0, z);
    }
}

void launch_residual(hipStream_t st, const double* A, int ld, int n, double alpha, const double* x,
                     const double* b, double* r) {
    residual_kernel<<<(n + 3) / 4, 256, 0, st>>>(A, ld, n, alpha, x, b, r);
}

void launch_add_diag(hipStream_t st, double* A, int ld, int n, double add) {
    add_diag_kernel<<<(n + 255) / 256, 256, 0, st>>>(A, ld, n, add);
}

void launch_axpy(hipStream_t st, int n, const double* d, double* x) {
    axpy_kernel<<<(n + 255) / 256, 256, 0, st>>>(n, d, x);
}

// padding granularity of the solve's matrices: whole 64-blocks
int solve_block() { return SB; }

}  // namespace plm
