// Residual pass over the weighted rows [X | y] of a sub-batch, for the
// iterative refinement of a ridge solution found from a quantised Gram and
// for exact squared-error sums:
//
//   u_k = y - X c_k                  (per row, FP64 sums)
//   R_k += X^T u_k,  s_k += u_k^T u_k (FP64)
//
// for up to eight coefficient vectors c_k (unscaled, zero past F). X is
// row-major with leading dimension ld and the weighted targets in column F;
// the coefficient rows are zero at column F and beyond, so the row products
// run over all ld columns without a bound check. Both kernels read X once
// (HBM bound); X^T u accumulates per workgroup over a row range and adds its
// column sums with one FP64 atomic per column and vector. ld is a multiple
// of 4 and X and C are 32-byte aligned (vector loads of four elements).
#include <hip/hip_runtime.h>

#include <stdexcept>

#include "kernels.hpp"

namespace plm {

constexpr int kResidRows = 4;    // rows per wave in the row products
constexpr int kXtrRows = 1024;   // rows per workgroup in X^T u

template <class T>
struct Vec4;
template <>
struct Vec4<float> {
    typedef float4 type;
};
template <>
struct Vec4<double> {
    typedef double4 type;
};

// one wave per kResidRows rows; lane l reads columns 4 (l + 64 t) .. + 3
template <class T, int N>
__global__ void __launch_bounds__(256) resid_rows_kernel(const T* __restrict__ X, int ld,
                                                         int rows, int F,
                                                         const double* __restrict__ C,
                                                         double* __restrict__ U) {
    typedef typename Vec4<T>::type V4;
    const int lane = threadIdx.x & 63;
    const int r0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * kResidRows;
    if (r0 >= rows) return;
    double acc[kResidRows][N];
#pragma unroll
    for (int q = 0; q < kResidRows; ++q)
#pragma unroll
        for (int k = 0; k < N; ++k) acc[q][k] = 0.0;
    const int nq = min(kResidRows, rows - r0);
    for (int c = 4 * lane; c < ld; c += 256) {
        double cv[N][4];
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const double2 a = *reinterpret_cast<const double2*>(C + (size_t)k * ld + c);
            const double2 b = *reinterpret_cast<const double2*>(C + (size_t)k * ld + c + 2);
            cv[k][0] = a.x, cv[k][1] = a.y, cv[k][2] = b.x, cv[k][3] = b.y;
        }
#pragma unroll
        for (int q = 0; q < kResidRows; ++q) {
            if (q >= nq) break;
            const V4 x = *reinterpret_cast<const V4*>(X + (size_t)(r0 + q) * ld + c);
            const double x0 = double(x.x), x1 = double(x.y), x2 = double(x.z), x3 = double(x.w);
#pragma unroll
            for (int k = 0; k < N; ++k)
                acc[q][k] = fma(x3, cv[k][3], fma(x2, cv[k][2], fma(x1, cv[k][1], fma(x0, cv[k][0], acc[q][k]))));
        }
    }
#pragma unroll
    for (int q = 0; q < kResidRows; ++q)
#pragma unroll
        for (int k = 0; k < N; ++k) {
            double v = acc[q][k];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
            acc[q][k] = v;
        }
    if (lane < nq) {
#pragma unroll
        for (int q = 0; q < kResidRows; ++q) {
            if (q != lane) continue;
            const double y = double(X[(size_t)(r0 + q) * ld + F]);
#pragma unroll
            for (int k = 0; k < N; ++k) U[(size_t)(r0 + q) * N + k] = y - acc[q][k];
        }
    }
}

// R[k][c] += sum_r X[r][c] U[r][k] over the workgroup's rows, one thread per
// column; the workgroups of column block 0 also add sum_r U[r][k]^2 to S[k].
template <class T, int N>
__global__ void __launch_bounds__(256) xtr_kernel(const T* __restrict__ X, int ld, int rows,
                                                  const double* __restrict__ U,
                                                  double* __restrict__ R, double* __restrict__ S) {
    __shared__ double red[4][N];
    const int c = blockIdx.x * 256 + threadIdx.x;
    const int cl = min(c, ld - 1);  // threads past ld read a valid column and add nothing
    const int r0 = blockIdx.y * kXtrRows, r1 = min(rows, r0 + kXtrRows);
    double acc[N];
#pragma unroll
    for (int k = 0; k < N; ++k) acc[k] = 0.0;
    // the row index and the U loads are uniform over the workgroup
    int r = r0;
    for (; r + 4 <= r1; r += 4) {
        double x[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) x[q] = double(X[(size_t)(r + q) * ld + cl]);
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int k = 0; k < N; ++k) acc[k] = fma(x[q], U[(size_t)(r + q) * N + k], acc[k]);
    }
    for (; r < r1; ++r) {
        const double x = double(X[(size_t)r * ld + cl]);
#pragma unroll
        for (int k = 0; k < N; ++k) acc[k] = fma(x, U[(size_t)r * N + k], acc[k]);
    }
    if (c < ld) {
#pragma unroll
        for (int k = 0; k < N; ++k) atomicAdd(R + (size_t)k * ld + c, acc[k]);
    }
    if (blockIdx.x != 0) return;
    double s[N];
#pragma unroll
    for (int k = 0; k < N; ++k) s[k] = 0.0;
    for (int q = r0 + threadIdx.x; q < r1; q += 256)
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const double u = U[(size_t)q * N + k];
            s[k] = fma(u, u, s[k]);
        }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        double v = s[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0) red[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < N)
        atomicAdd(S + threadIdx.x, red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] +
                                       red[3][threadIdx.x]);
}

template <class T, int N>
static void resid_launch(hipStream_t st, const void* X, int ld, int rows, int F, const double* C,
                         double* U, double* R, double* S) {
    const T* x = static_cast<const T*>(X);
    const int per_block = 4 * kResidRows;
    resid_rows_kernel<T, N><<<(rows + per_block - 1) / per_block, 256, 0, st>>>(x, ld, rows, F, C, U);
    const dim3 g((ld + 255) / 256, (rows + kXtrRows - 1) / kXtrRows);
    xtr_kernel<T, N><<<g, 256, 0, st>>>(x, ld, rows, U, R, S);
}

template <class T>
static void resid_dispatch(hipStream_t st, const void* X, int ld, int rows, int F, const double* C,
                           int n, double* U, double* R, double* S) {
    switch (n) {
        case 1: resid_launch<T, 1>(st, X, ld, rows, F, C, U, R, S); break;
        case 2: resid_launch<T, 2>(st, X, ld, rows, F, C, U, R, S); break;
        case 3: resid_launch<T, 3>(st, X, ld, rows, F, C, U, R, S); break;
        case 4: resid_launch<T, 4>(st, X, ld, rows, F, C, U, R, S); break;
        case 5: resid_launch<T, 5>(st, X, ld, rows, F, C, U, R, S); break;
        case 6: resid_launch<T, 6>(st, X, ld, rows, F, C, U, R, S); break;
        case 7: resid_launch<T, 7>(st, X, ld, rows, F, C, U, R, S); break;
        case 8: resid_launch<T, 8>(st, X, ld, rows, F, C, U, R, S); break;
        default: throw std::invalid_argument("residual pass: 1 to 8 coefficient vectors");
    }
}

void launch_residual_pass(hipStream_t st, bool fp64, const void* X, int ld, int rows, int F,
                          const double* C, int n, double* U, double* R, double* S) {
    if (ld % 4 || F >= ld || rows <= 0)
        throw std::invalid_argument("residual pass: leading dimension");
    if (fp64)
        resid_dispatch<double>(st, X, ld, rows, F, C, n, U, R, S);
    else
        resid_dispatch<float>(st, X, ld, rows, F, C, n, U, R, S);
}

}  // namespace plm
