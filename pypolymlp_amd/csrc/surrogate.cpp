// Host side of the surrogate network: device buffers and launches.
#include <vector>

#include "capi_internal.hpp"

namespace plm {

void surrogate_run(float* params, float* adam_m, float* adam_v, int t0, const float* x,
                   const float* y, int n, int steps, float lr, float wd, int mode, float* grad,
                   float* loss) {
    Device& dev = device();
    hipStream_t st = dev.stream;
    const size_t np = kSurrParams;
    DevBuf P, M, V, X, Y, G, L;
    P.alloc(np * 4), M.alloc(np * 4), V.alloc(np * 4), G.alloc(np * 4), L.alloc(4);
    X.alloc((size_t)n * kSurrIn * 4), Y.alloc((size_t)n * kSurrOut * 4);
    check_hip(hipMemcpyAsync(P.ptr, params, np * 4, hipMemcpyHostToDevice, st), "copy");
    if (mode == 0) {
        check_hip(hipMemcpyAsync(M.ptr, adam_m, np * 4, hipMemcpyHostToDevice, st), "copy");
        check_hip(hipMemcpyAsync(V.ptr, adam_v, np * 4, hipMemcpyHostToDevice, st), "copy");
    }
    check_hip(hipMemcpyAsync(X.ptr, x, (size_t)n * kSurrIn * 4, hipMemcpyHostToDevice, st), "copy");
    check_hip(hipMemcpyAsync(Y.ptr, y, (size_t)n * kSurrOut * 4, hipMemcpyHostToDevice, st), "copy");
    launch_surr_train(st, (float*)P.ptr, (float*)M.ptr, (float*)V.ptr, (const float*)X.ptr,
                      (const float*)Y.ptr, n, steps, t0, lr, 0.9f, 0.999f, 1e-8f, wd, mode,
                      (float*)G.ptr, (float*)L.ptr);
    check_hip(hipGetLastError(), "surrogate launch");
    if (mode == 0) {
        check_hip(hipMemcpyAsync(params, P.ptr, np * 4, hipMemcpyDeviceToHost, st), "copy");
        check_hip(hipMemcpyAsync(adam_m, M.ptr, np * 4, hipMemcpyDeviceToHost, st), "copy");
        check_hip(hipMemcpyAsync(adam_v, V.ptr, np * 4, hipMemcpyDeviceToHost, st), "copy");
    } else {
        check_hip(hipMemcpyAsync(grad, G.ptr, np * 4, hipMemcpyDeviceToHost, st), "copy");
    }
    check_hip(hipMemcpyAsync(loss, L.ptr, 4, hipMemcpyDeviceToHost, st), "copy");
    check_hip(hipStreamSynchronize(st), "sync");
}

void surrogate_predict(const float* params, const float* x, int n, float* y) {
    Device& dev = device();
    hipStream_t st = dev.stream;
    DevBuf P, X, Y;
    P.alloc((size_t)kSurrParams * 4);
    check_hip(hipMemcpyAsync(P.ptr, params, (size_t)kSurrParams * 4, hipMemcpyHostToDevice, st), "copy");
    for (int s0 = 0; s0 < n; s0 += kSurrMaxBatch) {
        const int m = std::min(kSurrMaxBatch, n - s0);
        X.alloc((size_t)m * kSurrIn * 4), Y.alloc((size_t)m * kSurrOut * 4);
        check_hip(hipMemcpyAsync(X.ptr, x + (size_t)s0 * kSurrIn, (size_t)m * kSurrIn * 4,
                                 hipMemcpyHostToDevice, st), "copy");
        launch_surr_predict(st, (const float*)P.ptr, (const float*)X.ptr, m, (float*)Y.ptr);
        check_hip(hipGetLastError(), "surrogate launch");
        check_hip(hipMemcpyAsync(y + (size_t)s0 * kSurrOut, Y.ptr, (size_t)m * kSurrOut * 4,
                                 hipMemcpyDeviceToHost, st), "copy");
        check_hip(hipStreamSynchronize(st), "sync");
    }
}

}  // namespace plm
