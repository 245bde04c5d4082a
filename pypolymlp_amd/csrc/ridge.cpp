// Ridge regression on the normal equations (solver_ridge,
// src/pypolymlp/mlp_dev/standard/solvers.py:48-84) on the GPU in FP64:
// Cholesky of (X^T X + alpha I), triangular solves and one step of iterative
// refinement with the residual of the unfactored system.
#include <algorithm>
#include <cstring>
#include <memory>
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

void launch_gram_system(hipStream_t st, const double* G, int ldg, int F, const double* s,
                        const unsigned char* zero, double* A, int lda, int n_pad, double* b);
void launch_symv(hipStream_t st, const double* A, int lda, int n, const double* x, double* y);

// For each alpha: Cholesky of (A0 + alpha I) (A0: full symmetric n_pad x
// n_pad on the device, identity on the padding), the solve with b, one
// refinement step with the residual of the unfactored system. coeffs: n_alpha
// x F on the host. b: one right-hand side of np entries on the device, or,
// with b_stride np, one per alpha.
// cache: factors of successful alphas are kept in it while it has room
// (kFactorBudget) and, for alphas it already holds, reused instead of
// factoring again.
static constexpr size_t kFactorBudget = 32ull << 30;

static void ridge_core(int F, int np, const DevBuf& A0, const DevBuf& b, const double* alphas,
                       int n_alpha, double* coeffs, int* info, size_t b_stride = 0,
                       FactorCache* cache = nullptr) {
    hipStream_t st = device().stream;
    const size_t mat = (size_t)np * np * 8;
    auto Af = std::make_unique<DevBuf>();
    DevBuf x, r, di;
    x.alloc(np * 8), r.alloc(np * 8), di.alloc(4);
    for (int k = 0; k < n_alpha; ++k) {
        const double a = alphas[k];
        const DevBuf* L = nullptr;
        if (cache)
            for (const auto& e : cache->f)
                if (e.first == a) L = e.second.get();
        int hi = 0;
        if (!L) {
            Af->alloc(mat);
            check_hip(hipMemcpyAsync(Af->ptr, A0.ptr, mat, hipMemcpyDeviceToDevice, st), "copy");
            launch_add_diag(st, (double*)Af->ptr, np, F, a);
            check_hip(hipMemsetAsync(di.ptr, 0, 4, st), "memset");
            {
                ProfScope ps("cholesky", st);
                launch_cholesky(st, (double*)Af->ptr, np, np, (int*)di.ptr);
            }
            check_hip(hipGetLastError(), "cholesky launch");
            check_hip(hipMemcpyAsync(&hi, di.ptr, 4, hipMemcpyDeviceToHost, st), "copy");
            check_hip(hipStreamSynchronize(st), "sync");
            L = Af.get();
        }
        {
            ProfScope ps("chol_solve", st);
            const double* bk = (const double*)b.ptr + k * b_stride;
            check_hip(hipMemcpyAsync(x.ptr, bk, np * 8, hipMemcpyDeviceToDevice, st), "copy");
            launch_chol_solve(st, (const double*)L->ptr, np, np, (double*)x.ptr);
            // one step of iterative refinement
            check_hip(hipMemsetAsync(r.ptr, 0, np * 8, st), "memset");
            launch_residual(st, (double*)A0.ptr, np, F, a, (double*)x.ptr, bk, (double*)r.ptr);
            launch_chol_solve(st, (const double*)L->ptr, np, np, (double*)r.ptr);
            launch_axpy(st, F, (double*)r.ptr, (double*)x.ptr);
        }
        check_hip(hipGetLastError(), "solve launch");
        check_hip(hipMemcpyAsync(coeffs + (size_t)k * F, x.ptr, (size_t)F * 8, hipMemcpyDeviceToHost, st), "copy");
        check_hip(hipStreamSynchronize(st), "sync");
        info[k] = hi;
        if (cache && L == Af.get() && hi == 0 && cache->bytes + mat <= kFactorBudget) {
            cache->bytes += mat;
            cache->f.emplace_back(a, std::move(Af));
            Af = std::make_unique<DevBuf>();
        }
    }
}

void ridge_solve(int F, const double* xtx, const double* xty, const double* alphas, int n_alpha,
                 double* coeffs, int* info) {
    if (F <= 0 || n_alpha <= 0) throw std::invalid_argument("empty ridge problem");
    const int B = solve_block();
    const int np = (F + B - 1) / B * B;
    const size_t mat = (size_t)np * np * 8;
    std::vector<double> h((size_t)np * np, 0.0);
    for (int i = 0; i < F; ++i) std::memcpy(&h[(size_t)i * np], xtx + (size_t)i * F, (size_t)F * 8);
    for (int i = F; i < np; ++i) h[(size_t)i * np + i] = 1.0;
    std::vector<double> hb(np, 0.0);
    std::memcpy(hb.data(), xty, (size_t)F * 8);
    DevBuf A0, b;
    A0.alloc(mat), b.alloc(np * 8);
    check_hip(hipMemcpy(A0.ptr, h.data(), mat, hipMemcpyHostToDevice), "copy");
    check_hip(hipMemcpy(b.ptr, hb.data(), np * 8, hipMemcpyHostToDevice), "copy");
    ridge_core(F, np, A0, b, alphas, n_alpha, coeffs, info);
}

// The scaled system of an accumulator, built on the device: A0 (np x np),
// b (np), and y^T y / b on the host for the squared-error sums.
static int build_system(NormalHandle& acc, const double* scales, const unsigned char* zero,
                        DevBuf& A0, DevBuf& b, std::vector<double>& hb, double& yty) {
    const int F = acc.F, B = solve_block();
    const int np = (F + B - 1) / B * B;
    hipStream_t st = device().stream;
    DevBuf ds, dz;
    ds.alloc((size_t)F * 8), dz.alloc((size_t)F);
    std::vector<unsigned char> z(F, 0);
    if (zero) std::memcpy(z.data(), zero, F);
    for (int i = 0; i < F; ++i)
        if (!z[i] && !(scales[i] != 0.0)) throw std::invalid_argument("zero scale on a kept column");
    check_hip(hipMemcpyAsync(ds.ptr, scales, (size_t)F * 8, hipMemcpyHostToDevice, st), "copy");
    check_hip(hipMemcpyAsync(dz.ptr, z.data(), (size_t)F, hipMemcpyHostToDevice, st), "copy");
    A0.alloc((size_t)np * np * 8), b.alloc((size_t)np * 8);
    launch_gram_system(st, (const double*)acc.xtx.ptr, acc.F_pad, F, (const double*)ds.ptr,
                       (const unsigned char*)dz.ptr, (double*)A0.ptr, np, np, (double*)b.ptr);
    check_hip(hipGetLastError(), "system launch");
    hb.assign(np, 0.0);
    check_hip(hipMemcpyAsync(hb.data(), b.ptr, (size_t)np * 8, hipMemcpyDeviceToHost, st), "copy");
    check_hip(hipMemcpyAsync(&yty, (const double*)acc.xtx.ptr + (size_t)F * acc.F_pad + F, 8,
                             hipMemcpyDeviceToHost, st), "copy");
    check_hip(hipStreamSynchronize(st), "sync");
    return np;
}

// c^T A0 c - 2 c^T b + y^T y for each coefficient vector (rows of coeffs)
static void sq_errors(int F, int np, const DevBuf& A0, const std::vector<double>& hb, double yty,
                      const double* coeffs, int n_coef, double* out) {
    hipStream_t st = device().stream;
    DevBuf dc, dv;
    dc.alloc((size_t)np * 8), dv.alloc((size_t)np * 8);
    std::vector<double> c(np, 0.0), v(np);
    for (int k = 0; k < n_coef; ++k) {
        std::memcpy(c.data(), coeffs + (size_t)k * F, (size_t)F * 8);
        check_hip(hipMemcpyAsync(dc.ptr, c.data(), (size_t)np * 8, hipMemcpyHostToDevice, st), "copy");
        launch_symv(st, (const double*)A0.ptr, np, np, (const double*)dc.ptr, (double*)dv.ptr);
        check_hip(hipGetLastError(), "symv launch");
        check_hip(hipMemcpyAsync(v.data(), dv.ptr, (size_t)np * 8, hipMemcpyDeviceToHost, st), "copy");
        check_hip(hipStreamSynchronize(st), "sync");
        double cav = 0.0, cb = 0.0;
        for (int i = 0; i < F; ++i) cav += c[i] * v[i], cb += c[i] * hb[i];
        out[k] = cav - 2.0 * cb + yty;
    }
}

void normal_ridge(NormalHandle& acc, const double* scales, const unsigned char* zero,
                  const double* alphas, int n_alpha, double* coeffs, int* info, double* sq_err) {
    if (n_alpha <= 0) throw std::invalid_argument("no alpha");
    check_hip(hipStreamSynchronize(device().stream), "sync");
    DevBuf A0, b;
    std::vector<double> hb;
    double yty = 0.0;
    const int np = build_system(acc, scales, zero, A0, b, hb, yty);
    // the factors are kept for a later correction step on the same system
    FactorCache& fc = acc.factors;
    fc.clear();
    fc.n_rows = acc.n_rows;
    fc.scales.assign(scales, scales + acc.F);
    fc.zero.assign(acc.F, 0);
    if (zero) fc.zero.assign(zero, zero + acc.F);
    ridge_core(acc.F, np, A0, b, alphas, n_alpha, coeffs, info, 0, &fc);
    if (sq_err) sq_errors(acc.F, np, A0, hb, yty, coeffs, n_alpha, sq_err);
}

void normal_sq_error(NormalHandle& acc, const double* scales, const unsigned char* zero,
                     const double* coeffs, int n_coef, double* sq_err) {
    check_hip(hipStreamSynchronize(device().stream), "sync");
    DevBuf A0, b;
    std::vector<double> hb;
    double yty = 0.0;
    const int np = build_system(acc, scales, zero, A0, b, hb, yty);
    sq_errors(acc.F, np, A0, hb, yty, coeffs, n_coef, sq_err);
}

// The correction step of iterative refinement for a solution c_k of the
// scaled system of an accumulator: delta_k = (A0 + alpha_k I)^{-1} rhs_k, with
// rhs_k = D^{-1} X^T (y - X D^{-1} c_k) - alpha_k c_k from a residual pass over
// the data (the residual of the normal equations of X, not of the
// accumulated Gram). rhs / delta: n_alpha x F on the host (dropped columns
// zero); dad[k] = delta_k^T A0 delta_k when dad is not null.
void normal_ridge_correct(NormalHandle& acc, const double* scales, const unsigned char* zero,
                          const double* alphas, int n_alpha, const double* rhs, double* delta,
                          int* info, double* dad) {
    if (n_alpha <= 0) throw std::invalid_argument("no alpha");
    hipStream_t st = device().stream;
    check_hip(hipStreamSynchronize(st), "sync");
    DevBuf A0, b;
    std::vector<double> hb;
    double yty = 0.0;
    const int F = acc.F;
    const int np = build_system(acc, scales, zero, A0, b, hb, yty);
    std::vector<unsigned char> z(F, 0);
    if (zero) std::memcpy(z.data(), zero, F);
    std::vector<double> hr((size_t)n_alpha * np, 0.0);
    for (int k = 0; k < n_alpha; ++k)
        for (int i = 0; i < F; ++i)
            if (!z[i]) hr[(size_t)k * np + i] = rhs[(size_t)k * F + i];
    DevBuf R;
    R.alloc(hr.size() * 8);
    check_hip(hipMemcpyAsync(R.ptr, hr.data(), hr.size() * 8, hipMemcpyHostToDevice, st), "copy");
    // the factors plm_normal_ridge kept, when they belong to this state and scaling
    FactorCache& fc = acc.factors;
    const bool same = fc.n_rows == acc.n_rows && fc.scales.size() == (size_t)F &&
                      std::equal(fc.scales.begin(), fc.scales.end(), scales) &&
                      std::equal(fc.zero.begin(), fc.zero.end(), z.begin());
    ridge_core(F, np, A0, R, alphas, n_alpha, delta, info, (size_t)np, same ? &fc : nullptr);
    if (dad) {
        const std::vector<double> zb(np, 0.0);
        sq_errors(F, np, A0, zb, 0.0, delta, n_alpha, dad);
    }
}

}  // namespace plm
