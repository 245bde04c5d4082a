// Conjugate gradient on (X^T X + alpha I) c = X^T y, device resident
// (solver_cg, src/pypolymlp/mlp_dev/gradient/solvers_cg.py:32-93).
//
// One iteration is three launches on one stream: symv (t = A d + alpha d and
// t.d), update (c += lr d, r -= lr t and |r|^2 with lr = |r|^2 / t.d) and
// direction (d = r + beta d). The scalars live in device memory, so the host
// only launches; the stopping test (iteration > 100 and |r| / sqrt(F) < gtol,
// made at the top of an iteration as the reference makes it) sets a flag that
// turns the remaining launches into no-ops.
//
// State S: S[0], S[1] |r|^2 of even / odd iterations, S[2] t.d, S[3] stop
// flag, S[4] iteration at which the test stopped the loop.
#include <hip/hip_runtime.h>

#include <cmath>

#include "kernels.hpp"

namespace plm {

constexpr int CG_THREADS = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// sum of v over the block, added to *dst by one atomic per block
__device__ __forceinline__ void block_sum_to(double v, double* dst) {
    __shared__ double part[CG_THREADS / 64];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) part[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < CG_THREADS / 64; ++w) s += part[w];
        atomicAdd(dst, s);
    }
}

// one wave per row of the symmetric A (row-major, leading dimension F)
__global__ void __launch_bounds__(CG_THREADS) cg_symv_kernel(const double* __restrict__ A, int F,
                                                              const double* __restrict__ d,
                                                              double alpha, double* __restrict__ t,
                                                              double* __restrict__ S, int it,
                                                              double gtol) {
    if (it >= 0) {
        if (S[3] != 0.0) return;
        const double rr = S[it & 1];
        if (it > 100 && sqrt(rr) / sqrt((double)F) < gtol) {
            if (blockIdx.x == 0 && threadIdx.x == 0) S[3] = 1.0, S[4] = it;
            return;
        }
        // accumulator of this iteration's new |r|^2 (last read two launches ago)
        if (blockIdx.x == 0 && threadIdx.x == 0) S[(it + 1) & 1] = 0.0;
    }
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * (CG_THREADS / 64) + (threadIdx.x >> 6);
    double td = 0.0;
    if (row < F) {
        const double* a = A + (size_t)row * F;
        double s = 0.0;
        for (int j = lane; j < F; j += 64) s += a[j] * d[j];
        s = wave_sum(s);
        if (lane == 0) {
            const double v = s + alpha * d[row];
            t[row] = v;
            td = v * d[row];
        }
    }
    block_sum_to(td, S + 2);
}

// r = b - t, d = r, |r|^2 -> S[0] (after cg_symv_kernel with it < 0 on d = c0)
__global__ void __launch_bounds__(CG_THREADS) cg_init_kernel(const double* __restrict__ b,
                                                              const double* __restrict__ t, int F,
                                                              double* __restrict__ r,
                                                              double* __restrict__ d,
                                                              double* __restrict__ S) {
    double rr = 0.0;
    for (int i = blockIdx.x * CG_THREADS + threadIdx.x; i < F; i += gridDim.x * CG_THREADS) {
        const double v = b[i] - t[i];
        r[i] = v, d[i] = v;
        rr += v * v;
    }
    // t.d of iteration 0 accumulates from zero (the symv on c0 wrote it)
    if (blockIdx.x == 0 && threadIdx.x == 0) S[2] = 0.0;
    block_sum_to(rr, S + 0);
}

__global__ void __launch_bounds__(CG_THREADS) cg_update_kernel(double* __restrict__ c,
                                                                double* __restrict__ r,
                                                                const double* __restrict__ d,
                                                                const double* __restrict__ t, int F,
                                                                double* __restrict__ S, int it) {
    if (S[3] != 0.0) return;
    const double lr = S[it & 1] / S[2];
    double rr = 0.0;
    for (int i = blockIdx.x * CG_THREADS + threadIdx.x; i < F; i += gridDim.x * CG_THREADS) {
        c[i] += lr * d[i];
        const double v = r[i] - lr * t[i];
        r[i] = v;
        rr += v * v;
    }
    block_sum_to(rr, S + ((it + 1) & 1));
}

__global__ void __launch_bounds__(CG_THREADS) cg_direction_kernel(double* __restrict__ d,
                                                                   const double* __restrict__ r,
                                                                   int F, double* __restrict__ S,
                                                                   int it) {
    if (S[3] != 0.0) return;
    const double beta = S[(it + 1) & 1] / S[it & 1];
    for (int i = blockIdx.x * CG_THREADS + threadIdx.x; i < F; i += gridDim.x * CG_THREADS)
        d[i] = r[i] + beta * d[i];
    // t.d of the next iteration accumulates from zero (update has read it)
    if (blockIdx.x == 0 && threadIdx.x == 0) S[2] = 0.0;
}

static int vec_grid(int F) {
    const int g = (F + CG_THREADS - 1) / CG_THREADS;
    return g < 1024 ? g : 1024;
}

void launch_cg_symv(hipStream_t st, const double* A, int F, const double* d, double alpha,
                    double* t, double* S, int it, double gtol) {
    const int rows_per_block = CG_THREADS / 64;
    cg_symv_kernel<<<(F + rows_per_block - 1) / rows_per_block, CG_THREADS, 0, st>>>(
        A, F, d, alpha, t, S, it, gtol);
}

void launch_cg_init(hipStream_t st, const double* b, const double* t, int F, double* r, double* d,
                    double* S) {
    cg_init_kernel<<<vec_grid(F), CG_THREADS, 0, st>>>(b, t, F, r, d, S);
}

void launch_cg_update(hipStream_t st, double* c, double* r, const double* d, const double* t,
                      int F, double* S, int it) {
    cg_update_kernel<<<vec_grid(F), CG_THREADS, 0, st>>>(c, r, d, t, F, S, it);
}

void launch_cg_direction(hipStream_t st, double* d, const double* r, int F, double* S, int it) {
    cg_direction_kernel<<<vec_grid(F), CG_THREADS, 0, st>>>(d, r, F, S, it);
}

}  // namespace plm
