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

// Unblocked Cholesky of the diagonal block at (k0, k0) in LDS.
__global__ void __launch_bounds__(256) potf2_kernel(double* __restrict__ A, int ld, int k0,
                                                    int* __restrict__ info) {
    __shared__ double a[SB][SB + 1];
    const int tid = threadIdx.x;
    for (int e = tid; e < SB * SB; e += 256) {
        const int i = e / SB, j = e % SB;
        a[i][j] = (j <= i) ? A[(size_t)(k0 + i) * ld + k0 + j] : 0.0;
    }
    __syncthreads();
    for (int j = 0; j < SB; ++j) {
        if (tid == 0) {
            double d = a[j][j];
            if (!(d > 0.0)) {
                if (*info == 0) *info = k0 + j + 1;
                d = 1.0;  // keep the factor finite; the caller discards this solve
            }
            a[j][j] = sqrt(d);
        }
        __syncthreads();
        const double djj = a[j][j];
        for (int i = j + 1 + tid; i < SB; i += 256) a[i][j] /= djj;
        __syncthreads();
        const int m = SB - j - 1;
        for (int e = tid; e < m * m; e += 256) {
            const int i = j + 1 + e / m, c = j + 1 + e % m;
            if (c <= i) a[i][c] -= a[i][j] * a[c][j];
        }
        __syncthreads();
    }
    for (int e = tid; e < SB * SB; e += 256) {
        const int i = e / SB, j = e % SB;
        if (j <= i) A[(size_t)(k0 + i) * ld + k0 + j] = a[i][j];
    }
}

// Panel: rows i in [k0+SB, n): A[i, k0:k0+SB] <- A[i, k0:k0+SB] L_kk^{-T}.
// One thread per row, the diagonal block in LDS.
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
    for (int c = 0; c < SB; ++c) x[c] = row[c];
    for (int c = 0; c < SB; ++c) {
        double v = x[c];
        for (int cp = 0; cp < c; ++cp) v -= x[cp] * L[c][cp];
        x[c] = v / L[c][c];
    }
    for (int c = 0; c < SB; ++c) row[c] = x[c];
}

// Trailing update of the lower triangle: for 64x64 tiles (bi >= bj) of the
// trailing matrix, C -= P_bi P_bj^T with the panel P (K = SB) on FP64 MFMA.
__device__ __forceinline__ void low_tri_index(int t, int& bi, int& bj) {
    // row bi holds bi + 1 tiles
    int r = 0, rem = t;
    while (rem >= r + 1) rem -= r + 1, ++r;
    bi = r, bj = rem;
}

__global__ void __launch_bounds__(256) syrk_update_kernel(double* __restrict__ A, int ld, int k0) {
    __shared__ double Ps[SB][SB + 1];  // [k][i] panel rows of the I tile
    __shared__ double Qs[SB][SB + 1];  // [k][j] panel rows of the J tile
    int bi, bj;
    low_tri_index(blockIdx.x, bi, bj);
    const int base = k0 + SB;
    const int I0 = base + bi * SB, J0 = base + bj * SB;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int e = tid; e < SB * SB; e += 256) {
        const int r = e / SB, k = e % SB;
        Ps[k][r] = A[(size_t)(I0 + r) * ld + k0 + k];
        Qs[k][r] = A[(size_t)(J0 + r) * ld + k0 + k];
    }
    __syncthreads();
    const int wi = (wave >> 1) * 32, wj = (wave & 1) * 32;
    doublex4 acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = doublex4{0, 0, 0, 0};
    for (int ks = 0; ks < SB; ks += 4) {
        const int kk = ks + (lane >> 4);
        double av[2], bv[2];
        for (int m = 0; m < 2; ++m) av[m] = Ps[kk][wi + 16 * m + (lane & 15)];
        for (int q = 0; q < 2; ++q) bv[q] = Qs[kk][wj + 16 * q + (lane & 15)];
        for (int m = 0; m < 2; ++m)
            for (int q = 0; q < 2; ++q)
                acc[m][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[m], bv[q], acc[m][q], 0, 0, 0);
    }
    for (int m = 0; m < 2; ++m)
        for (int q = 0; q < 2; ++q)
            for (int v = 0; v < 4; ++v) {
                const int i = I0 + wi + 16 * m + (lane >> 4) + 4 * v;
                const int j = J0 + wj + 16 * q + (lane & 15);
                if (j <= i) A[(size_t)i * ld + j] -= acc[m][q][v];
            }
}

// Forward substitution on one diagonal block: z_k = L_kk^{-1} z_k.
__global__ void __launch_bounds__(64) trsv_lower_block(const double* __restrict__ L, int ld, int k0,
                                                       double* __restrict__ z) {
    __shared__ double Lb[SB][SB + 1];
    __shared__ double zb[SB];
    for (int e = threadIdx.x; e < SB * SB; e += 64) {
        const int i = e / SB, j = e % SB;
        Lb[i][j] = (j <= i) ? L[(size_t)(k0 + i) * ld + k0 + j] : 0.0;
    }
    zb[threadIdx.x] = z[k0 + threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int c = 0; c < SB; ++c) {
            double v = zb[c];
            for (int cp = 0; cp < c; ++cp) v -= Lb[c][cp] * zb[cp];
            zb[c] = v / Lb[c][c];
        }
    }
    __syncthreads();
    z[k0 + threadIdx.x] = zb[threadIdx.x];
}

// Backward substitution on one diagonal block: x_k = L_kk^{-T} x_k.
__global__ void __launch_bounds__(64) trsv_upper_block(const double* __restrict__ L, int ld, int k0,
                                                       double* __restrict__ z) {
    __shared__ double Lb[SB][SB + 1];
    __shared__ double zb[SB];
    for (int e = threadIdx.x; e < SB * SB; e += 64) {
        const int i = e / SB, j = e % SB;
        Lb[i][j] = (j <= i) ? L[(size_t)(k0 + i) * ld + k0 + j] : 0.0;
    }
    zb[threadIdx.x] = z[k0 + threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int c = SB - 1; c >= 0; --c) {
            double v = zb[c];
            for (int cp = c + 1; cp < SB; ++cp) v -= Lb[cp][c] * zb[cp];
            zb[c] = v / Lb[c][c];
        }
    }
    __syncthreads();
    z[k0 + threadIdx.x] = zb[threadIdx.x];
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

// r = b - (A + alpha I) x with the full symmetric A (one thread per row,
// FP64 with a compensated (TwoSum-style) accumulation of the products).
__global__ void residual_kernel(const double* __restrict__ A, int ld, int n, double alpha,
                                const double* __restrict__ x, const double* __restrict__ b,
                                double* __restrict__ r) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* row = A + (size_t)i * ld;
    double s = 0.0, comp = 0.0;
    for (int c = 0; c < n; ++c) {
        const double p = row[c] * x[c];
        const double t = s + p;
        comp += (fabs(s) >= fabs(p)) ? (s - t) + p : (p - t) + s;
        s = t;
    }
    s += comp;
    r[i] = b[i] - (s + alpha * x[i]);
}

__global__ void add_diag_kernel(double* __restrict__ A, int ld, int n, double add) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) A[(size_t)i * ld + i] += add;
}

__global__ void axpy_kernel(int n, const double* __restrict__ d, double* __restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += d[i];
}

void launch_cholesky(hipStream_t st, double* A, int ld, int n_pad, int* info) {
    for (int k0 = 0; k0 < n_pad; k0 += SB) {
        potf2_kernel<<<1, 256, 0, st>>>(A, ld, k0, info);
        const int rest = n_pad - k0 - SB;
        if (rest <= 0) break;
        trsm_panel_kernel<<<(rest + 255) / 256, 256, 0, st>>>(A, ld, k0, n_pad);
        const int nt = rest / SB;
        syrk_update_kernel<<<nt * (nt + 1) / 2, 256, 0, st>>>(A, ld, k0);
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
    residual_kernel<<<(n + 255) / 256, 256, 0, st>>>(A, ld, n, alpha, x, b, r);
}

void launch_add_diag(hipStream_t st, double* A, int ld, int n, double add) {
    add_diag_kernel<<<(n + 255) / 256, 256, 0, st>>>(A, ld, n, add);
}

void launch_axpy(hipStream_t st, int n, const double* d, double* x) {
    axpy_kernel<<<(n + 255) / 256, 256, 0, st>>>(n, d, x);
}

int solve_block() { return SB; }

}  // namespace plm
