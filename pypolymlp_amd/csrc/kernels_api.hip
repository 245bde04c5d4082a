// Device side of the libmlpcpp support functions: the radial basis and the
// spherical harmonics of given points (get_fn / get_ylm,
// src/cxx/src/python/pybind11_mlp.cpp:162-184) and the conjugate fill of a
// synthetic order-parameter vector (PolymlpAPI::compute_anlmtp_conjugate,
// src/cxx/src/polymlp/polymlp_api.cpp:63-87). They run the same device
// functions (device_math.hpp) and type tables as the X and evaluation kernels,
// so that the reference's known answers for these functions pin the device
// code directly.
#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "device_tables.hpp"
#include "kernels.hpp"

namespace plm {

// One thread per distance: f_n(r_k), f'_n(r_k) for the n_par (beta, mu) pairs.
template <class T>
__global__ void get_fn_kernel(int n, const double* __restrict__ r, const double* __restrict__ par,
                              int n_par, double cutoff, double* __restrict__ fn,
                              double* __restrict__ fd) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    T f[kMaxFn], d[kMaxFn];
    radial<T>(T(r[k]), T(cutoff), par, n_par, f, d);
    for (int i = 0; i < n_par; ++i) {
        fn[(size_t)k * n_par + i] = double(f[i]);
        fd[(size_t)k * n_par + i] = double(d[i]);
    }
}

// One thread per point: Y_l^{-m} at key l(l+3)/2 - m and dY/dx, dY/dy, dY/dz,
// written as (re, im) pairs: y_out[k][key][2], dy_out[k][key][a][2].
template <class T>
__global__ void get_ylm_kernel(int n, const double* __restrict__ xyz, int lmax,
                               double* __restrict__ y_out, double* __restrict__ dy_out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int nh = (lmax + 1) * (lmax + 2) / 2;
    Cpx<T> Y[kMaxLmHalf], dY[3 * kMaxLmHalf];
    T p[kMaxLmHalf], q[kMaxLmHalf];
    const double x = xyz[3 * k], y = xyz[3 * k + 1], z = xyz[3 * k + 2];
    const double r = sqrt(x * x + y * y + z * z);
    ylm<T>(lmax, T(x), T(y), T(z), T(r), Y, dY, p, q);
    for (int i = 0; i < nh; ++i) {
        y_out[((size_t)k * nh + i) * 2] = double(Y[i].re);
        y_out[((size_t)k * nh + i) * 2 + 1] = double(Y[i].im);
        for (int a = 0; a < 3; ++a) {
            dy_out[(((size_t)k * nh + i) * 3 + a) * 2] = double(dY[3 * i + a].re);
            dy_out[(((size_t)k * nh + i) * 3 + a) * 2 + 1] = double(dY[3 * i + a].im);
        }
    }
}

// a_full[k] = a_nc[src[k]] (plain) or sign[k] conj(a_nc[src[k]]) (conj[k])
template <class T>
__global__ void conj_fill_kernel(int n_nlm, const int* __restrict__ src,
                                 const double* __restrict__ sign,
                                 const unsigned char* __restrict__ conj,
                                 const double* __restrict__ re, const double* __restrict__ im,
                                 T* __restrict__ a_full) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_nlm) return;
    const int s = src[k];
    const double sg = sign[k];
    const bool cj = conj[k];
    a_full[2 * k] = T(cj ? sg * re[s] : re[s]);
    a_full[2 * k + 1] = T(cj ? -sg * im[s] : im[s]);
}

void launch_get_fn(hipStream_t st, bool fp64, int n, const double* r, const double* par, int n_par,
                   double cutoff, double* fn, double* fd) {
    const int B = 64;
    if (fp64)
        get_fn_kernel<double><<<(n + B - 1) / B, B, 0, st>>>(n, r, par, n_par, cutoff, fn, fd);
    else
        get_fn_kernel<float><<<(n + B - 1) / B, B, 0, st>>>(n, r, par, n_par, cutoff, fn, fd);
}

void launch_get_ylm(hipStream_t st, bool fp64, int n, const double* xyz, int lmax, double* ylm,
                    double* dylm) {
    const int B = 64;
    if (fp64)
        get_ylm_kernel<double><<<(n + B - 1) / B, B, 0, st>>>(n, xyz, lmax, ylm, dylm);
    else
        get_ylm_kernel<float><<<(n + B - 1) / B, B, 0, st>>>(n, xyz, lmax, ylm, dylm);
}

void launch_conj_fill(hipStream_t st, bool fp64, int n_nlm, const int* src, const double* sign,
                      const unsigned char* conj, const double* re, const double* im,
                      void* a_full) {
    const int B = 256, g = (n_nlm + B - 1) / B;
    if (g == 0) return;
    if (fp64)
        conj_fill_kernel<double><<<g, B, 0, st>>>(n_nlm, src, sign, conj, re, im,
                                                  static_cast<double*>(a_full));
    else
        conj_fill_kernel<float><<<g, B, 0, st>>>(n_nlm, src, sign, conj, re, im,
                                                 static_cast<float*>(a_full));
}

}  // namespace plm
