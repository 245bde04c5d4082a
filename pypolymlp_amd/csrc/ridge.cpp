// Ridge regression on the normal equations (solver_ridge,
// src/pypolymlp/mlp_dev/standard/solvers.py:48-84) on the GPU in FP64:
// Cholesky of (X^T X + alpha I), triangular solves and one step of iterative
// refinement with the residual of the unfactored system.
#include <cstring>
#include <vector>

#include "capi_internal.hpp"
#include "profile.hpp"

namespace plm {

void launch_cholesky(hipStream_t st, double* A, int ld, int n_pad, int* info);
void launch_chol_solve(hipStream_t st, const double* L, int ld, int n_pad, double* z);
void launch_residual(hipStream_t st, const double* A, int ld, int n, double alpha, const double* x,
                     const double* b, double* r);
void launch_add_diag(hipStream_t st, double* A, int ld, int n, double add);
void launch_axpy(hipStream_t st, int n, const double* d, double* x);
int solve_block();

void ridge_solve(int F, const double* xtx, const double* xty, const double* alphas, int n_alpha,
                 double* coeffs, int* info) {
    if (F <= 0 || n_alpha <= 0) throw std::invalid_argument("empty ridge problem");
    Device& dev = device();
    hipStream_t st = dev.stream;
    const int B = solve_block();
    const int np = (F + B - 1) / B * B;
    const size_t mat = (size_t)np * np * 8;
    std::vector<double> h((size_t)np * np, 0.0);
    for (int i = 0; i < F; ++i) std::memcpy(&h[(size_t)i * np], xtx + (size_t)i * F, (size_t)F * 8);
    for (int i = F; i < np; ++i) h[(size_t)i * np + i] = 1.0;
    std::vector<double> hb(np, 0.0);
    std::memcpy(hb.data(), xty, (size_t)F * 8);
    DevBuf A0, Af, b, x, r, di;
    A0.alloc(mat), Af.alloc(mat), b.alloc(np * 8), x.alloc(np * 8), r.alloc(np * 8), di.alloc(4);
    check_hip(hipMemcpy(A0.ptr, h.data(), mat, hipMemcpyHostToDevice), "copy");
    check_hip(hipMemcpy(b.ptr, hb.data(), np * 8, hipMemcpyHostToDevice), "copy");
    for (int k = 0; k < n_alpha; ++k) {
        const double a = alphas[k];
        check_hip(hipMemcpyAsync(Af.ptr, A0.ptr, mat, hipMemcpyDeviceToDevice, st), "copy");
        launch_add_diag(st, (double*)Af.ptr, np, F, a);
        check_hip(hipMemsetAsync(di.ptr, 0, 4, st), "memset");
        {
            ProfScope ps("cholesky", st);
            launch_cholesky(st, (double*)Af.ptr, np, np, (int*)di.ptr);
        }
        check_hip(hipGetLastError(), "cholesky launch");
        check_hip(hipMemcpyAsync(x.ptr, b.ptr, np * 8, hipMemcpyDeviceToDevice, st), "copy");
        launch_chol_solve(st, (double*)Af.ptr, np, np, (double*)x.ptr);
        // one step of iterative refinement
        check_hip(hipMemsetAsync(r.ptr, 0, np * 8, st), "memset");
        launch_residual(st, (double*)A0.ptr, np, F, a, (double*)x.ptr, (double*)b.ptr, (double*)r.ptr);
        launch_chol_solve(st, (double*)Af.ptr, np, np, (double*)r.ptr);
        launch_axpy(st, F, (double*)r.ptr, (double*)x.ptr);
        check_hip(hipGetLastError(), "solve launch");
        int hi = 0;
        check_hip(hipMemcpyAsync(coeffs + (size_t)k * F, x.ptr, (size_t)F * 8, hipMemcpyDeviceToHost, st), "copy");
        check_hip(hipMemcpyAsync(&hi, di.ptr, 4, hipMemcpyDeviceToHost, st), "copy");
        check_hip(hipStreamSynchronize(st), "sync");
        info[k] = hi;
    }
}

}  // namespace plm
