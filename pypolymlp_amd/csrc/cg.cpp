// Conjugate-gradient solve of the regularised normal equations on the GPU
// (solver_cg, src/pypolymlp/mlp_dev/gradient/solvers_cg.py:32-93).
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "capi_internal.hpp"

namespace plm {

void launch_cg_symv(hipStream_t st, const double* A, int F, const double* d, double alpha,
                    double* t, double* S, int it, double gtol);
void launch_cg_init(hipStream_t st, const double* b, const double* t, int F, double* r, double* d,
                    double* S);
void launch_cg_update(hipStream_t st, double* c, double* r, const double* d, const double* t,
                      int F, double* S, int it);
void launch_cg_direction(hipStream_t st, double* d, const double* r, int F, double* S, int it);

// Iterations are launched in chunks; after each chunk the host reads the stop
// flag, so a converged solve ends at most one chunk of no-op launches later.
constexpr int kCgChunk = 64;

int cg_solve(int F, const double* xtx, const double* xty, double alpha, const double* coef0,
             double gtol, int max_iter, double* coef) {
    if (F <= 0) throw std::invalid_argument("empty CG problem");
    if (max_iter < 0) throw std::invalid_argument("negative max_iter");
    Device& dev = device();
    hipStream_t st = dev.stream;
    DevBuf A, b, c, r, d, t, S;
    const size_t vb = (size_t)F * 8;
    A.alloc((size_t)F * F * 8), b.alloc(vb), c.alloc(vb), r.alloc(vb), d.alloc(vb), t.alloc(vb);
    S.alloc(8 * 8);
    check_hip(hipMemcpyAsync(A.ptr, xtx, (size_t)F * F * 8, hipMemcpyHostToDevice, st), "copy");
    check_hip(hipMemcpyAsync(b.ptr, xty, vb, hipMemcpyHostToDevice, st), "copy");
    if (coef0)
        check_hip(hipMemcpyAsync(c.ptr, coef0, vb, hipMemcpyHostToDevice, st), "copy");
    else
        check_hip(hipMemsetAsync(c.ptr, 0, vb, st), "memset");
    check_hip(hipMemsetAsync(S.ptr, 0, 8 * 8, st), "memset");
    double* Sd = (double*)S.ptr;
    // r = b - (A + alpha I) c0, d = r
    launch_cg_symv(st, (double*)A.ptr, F, (double*)c.ptr, alpha, (double*)t.ptr, Sd, -1, gtol);
    launch_cg_init(st, (double*)b.ptr, (double*)t.ptr, F, (double*)r.ptr, (double*)d.ptr, Sd);
    check_hip(hipGetLastError(), "cg init launch");
    double hs[8] = {0};
    for (int it0 = 0; it0 < max_iter; it0 += kCgChunk) {
        const int it1 = std::min(max_iter, it0 + kCgChunk);
        for (int it = it0; it < it1; ++it) {
            launch_cg_symv(st, (double*)A.ptr, F, (double*)d.ptr, alpha, (double*)t.ptr, Sd, it,
                           gtol);
            launch_cg_update(st, (double*)c.ptr, (double*)r.ptr, (double*)d.ptr,
                             (double*)t.ptr, F, Sd, it);
            launch_cg_direction(st, (double*)d.ptr, (double*)r.ptr, F, Sd, it);
        }
        check_hip(hipGetLastError(), "cg launch");
        check_hip(hipMemcpyAsync(hs, Sd, sizeof hs, hipMemcpyDeviceToHost, st), "copy");
        check_hip(hipStreamSynchronize(st), "sync");
        if (hs[3] != 0.0) break;
    }
    check_hip(hipMemcpyAsync(coef, c.ptr, vb, hipMemcpyDeviceToHost, st), "copy");
    check_hip(hipMemcpyAsync(hs, Sd, sizeof hs, hipMemcpyDeviceToHost, st), "copy");
    check_hip(hipStreamSynchronize(st), "sync");
    return hs[3] != 0.0 ? (int)hs[4] : max_iter;
}

}  // namespace plm
