// Device math for one neighbour pair: Gaussian x cosine-cutoff radial
// functions and complex real-normalised spherical harmonics Y_l^{-m} (m >= 0)
// with Cartesian derivatives.
//
// Radial: f_n(r) = exp(-beta (r - mu)^2) * fc(r), fc = (cos(pi r / rc) + 1) / 2
//   (polymlp_basis_function.h:16-43, polymlp_functions_interface.cpp:41-63).
// Angular: normalised associated Legendre functions by the recursion of
//   Limpanuparb & Milthorpe (arXiv:1410.1748) in the variant used by the
//   reference (polymlp_spherical_harmonics.cpp:27-207): P and Q = P / sin(theta),
//   Chebyshev recurrence for cos/sin(m phi), Y stored at key l(l+3)/2 - m.
#pragma once

#include <hip/hip_runtime.h>

namespace plm {

template <class T>
struct Cpx {
    T re, im;
};

template <class T>
__device__ __forceinline__ T re_mul(Cpx<T> a, Cpx<T> b) {
    return a.re * b.re - a.im * b.im;
}
template <class T>
__device__ __forceinline__ Cpx<T> cmul(Cpx<T> a, Cpx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// dst[k] = src[k] for k < n complex values, strided over the workgroup, eight
// loads in flight per thread (as the vector type of two T)
template <class T> struct Vec2Of;
template <> struct Vec2Of<float> { using type = float2; };
template <> struct Vec2Of<double> { using type = double2; };
template <class T> struct Vec4Of;
template <> struct Vec4Of<float> { using type = float4; };
template <> struct Vec4Of<double> { using type = double4; };
template <class T>
__device__ __forceinline__ void stage_cpx(Cpx<T>* dst, const Cpx<T>* __restrict__ src, int n) {
    using V = typename Vec2Of<T>::type;
    const V* s = reinterpret_cast<const V*>(src);
    V* d = reinterpret_cast<V*>(dst);
    for (int k = threadIdx.x; k < n; k += 8 * blockDim.x) {
        V v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int kk = k + u * blockDim.x;
            v[u] = kk < n ? s[kk] : V{};
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int kk = k + u * blockDim.x;
            if (kk < n) d[kk] = v[u];
        }
    }
}

template <class T> struct MathConst;
template <> struct MathConst<double> {
    static constexpr double pi = 3.1415926535897932384626433832795;
    static constexpr double sqrt_inv_2pi = 0.39894228040143267794;
    static constexpr double sqrt3 = 1.7320508075688772935;
    static constexpr double sqrt3div2 = 1.2247448713915890491;
    static constexpr double half_sqrt2 = 0.70710678118654752440;
};
template <> struct MathConst<float> {
    static constexpr float pi = 3.14159265358979323846f;
    static constexpr float sqrt_inv_2pi = 0.39894228040143267794f;
    static constexpr float sqrt3 = 1.7320508075688772935f;
    static constexpr float sqrt3div2 = 1.2247448713915890491f;
    static constexpr float half_sqrt2 = 0.70710678118654752440f;
};

__device__ __forceinline__ double dsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float dsqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double dhypot(double x, double y) { return hypot(x, y); }
__device__ __forceinline__ float dhypot(float x, float y) { return hypotf(x, y); }
__device__ __forceinline__ double dexp(double x) { return exp(x); }
__device__ __forceinline__ float dexp(float x) { return __expf(x); }
__device__ __forceinline__ void dsincos(double x, double* s, double* c) { sincos(x, s, c); }
__device__ __forceinline__ void dsincos(float x, float* s, float* c) { sincosf(x, s, c); }

// Radial functions f_n and df_n/dr for the n_id list of one type pair.
// rad: (beta, mu) pairs. Returns nothing for r >= rc (caller skips).
template <class T>
__device__ __forceinline__ void radial(T r, T rc, const double* rad, int count, T* fn, T* fd) {
    const T k = MathConst<T>::pi / rc;
    T s, c;
    dsincos(k * r, &s, &c);
    const T fc = T(0.5) * (c + T(1));
    const T fcd = T(-0.5) * k * s;
    for (int n = 0; n < count; ++n) {
        const T beta = T(rad[2 * n]), mu = T(rad[2 * n + 1]);
        const T dx = r - mu;
        const T g = dexp(-beta * dx * dx);
        const T gd = T(-2) * beta * dx * g;
        fn[n] = g * fc;
        fd[n] = gd * fc + g * fcd;
    }
}

template <class T>
__device__ __forceinline__ void radial_value(T r, T rc, const double* rad, int count, T* fn) {
    const T k = MathConst<T>::pi / rc;
    const T fc = T(0.5) * (cos(k * r) + T(1));
    for (int n = 0; n < count; ++n) {
        const T beta = T(rad[2 * n]), mu = T(rad[2 * n + 1]);
        const T dx = r - mu;
        fn[n] = dexp(-beta * dx * dx) * fc;
    }
}

__host__ __device__ constexpr int lm2i(int l, int m) { return l * (l + 1) / 2 + m; }

// Normalised associated Legendre P_l^m(cos t) (and Q = P / sin t when Q != nullptr)
// for 0 <= m <= l <= lmax, stored at lm2i(l, m).
template <class T>
__device__ __forceinline__ void legendre(int lmax, T ct, T st, T* p, T* q) {
    using C = MathConst<T>;
    p[0] = C::sqrt_inv_2pi;
    if (q) q[0] = T(0);
    if (lmax == 0) return;
    p[lm2i(1, 0)] = ct * C::sqrt3 * C::sqrt_inv_2pi;
    p[lm2i(1, 1)] = -st * C::sqrt3div2 * C::sqrt_inv_2pi;
    if (q) {
        q[lm2i(1, 0)] = T(0);
        q[lm2i(1, 1)] = -C::sqrt3div2 * C::sqrt_inv_2pi;
    }
    for (int l = 2; l <= lmax; ++l) {
        const T c1 = -dsqrt(T(1) + T(0.5) / T(l)) * st;
        p[lm2i(l, l)] = c1 * p[lm2i(l - 1, l - 1)];
        const T c2 = dsqrt(T(2) * (T(l) - T(1)) + T(3)) * ct;
        p[lm2i(l, l - 1)] = c2 * p[lm2i(l - 1, l - 1)];
        if (q) {
            q[lm2i(l, l)] = c1 * q[lm2i(l - 1, l - 1)];
            q[lm2i(l, l - 1)] = c2 * q[lm2i(l - 1, l - 1)];
        }
    }
    for (int l = 2; l <= lmax; ++l) {
        const T ls = T(l * l), lm1s = T((l - 1) * (l - 1));
        for (int m = 0; m <= l - 2; ++m) {
            const T ms = T(m * m);
            const T alm = dsqrt((T(4) * ls - T(1)) / (ls - ms));
            const T blm = -dsqrt((lm1s - ms) / (T(4) * lm1s - T(1)));
            p[lm2i(l, m)] = alm * (ct * p[lm2i(l - 1, m)] + blm * p[lm2i(l - 2, m)]);
            if (q) q[lm2i(l, m)] = alm * (ct * q[lm2i(l - 1, m)] + blm * q[lm2i(l - 2, m)]);
        }
    }
}

// Y_l^{-m} (m >= 0) at key l(l+3)/2 - m for the direction (x, y, z) of length r.
// If dy != nullptr, also dY/dx, dY/dy, dY/dz (dy[3*key + a]).
// p, q: scratch of size (lmax+1)(lmax+2)/2 each.
template <class T>
__device__ __forceinline__ void ylm(int lmax, T x, T y, T z, T r, Cpx<T>* Y, Cpx<T>* dY,
                                    T* p, T* q) {
    using C = MathConst<T>;
    const T ct = z / r;
    const T rho = dhypot(x, y);
    T cp = T(1), sp = T(0);
    if (rho > T(0)) cp = x / rho, sp = y / rho;
    const T st = dsqrt(T(1) - ct * ct);
    const T invr = T(1) / r;
    legendre<T>(lmax, ct, st, p, dY ? q : nullptr);
    for (int l = 0; l <= lmax; ++l) {
        const int key = lm2i(l, 0) + l;
        Y[key] = {p[lm2i(l, 0)] * C::half_sqrt2, T(0)};
        if (dY) {
            const T common =
                (l >= 1) ? q[lm2i(l, 1)] * st * invr * dsqrt(T(0.5) * T(l) * T(l + 1)) : T(0);
            dY[3 * key + 0] = {common * ct * cp, T(0)};
            dY[3 * key + 1] = {common * ct * sp, T(0)};
            dY[3 * key + 2] = {-common * st, T(0)};
        }
    }
    T c1 = T(1), c2 = cp, s1 = T(0), s2 = -sp;
    const T tc = T(2) * c2;
    T sign = T(-1);
    for (int mp = 1; mp <= lmax; ++mp) {
        const T s = tc * s1 - s2;
        const T c = tc * c1 - c2;
        c2 = c1, c1 = c, s2 = s1, s1 = s;
        for (int l = mp; l <= lmax; ++l) {
            const int key = lm2i(l, -mp) + l;
            const T tmp = sign * p[lm2i(l, mp)] * C::half_sqrt2;
            Y[key] = {tmp * c, -tmp * s};
            if (dY) {
                // common = e^{i m phi} / sqrt(2) / r
                const T cre = c * C::half_sqrt2 * invr, cim = s * C::half_sqrt2 * invr;
                T dth = T(mp) * ct * q[lm2i(l, mp)];
                if (mp != l) dth += dsqrt(T((l - mp) * (l + mp + 1))) * q[lm2i(l, mp + 1)] * st;
                const T dph = T(mp) * q[lm2i(l, mp)];  // imaginary
                // u = dth*ct*cp - i*dph*sp ; v = dth*ct*sp + i*dph*cp ; w = -dth*st
                const T ure = dth * ct * cp, uim = -dph * sp;
                const T vre = dth * ct * sp, vim = dph * cp;
                const T wre = -dth * st;
                // sign * conj(common * u)
                dY[3 * key + 0] = {sign * (cre * ure - cim * uim), -sign * (cre * uim + cim * ure)};
                dY[3 * key + 1] = {sign * (cre * vre - cim * vim), -sign * (cre * vim + cim * vre)};
                dY[3 * key + 2] = {sign * (cre * wre), -sign * (cim * wre)};
            }
        }
        sign = -sign;
    }
}

}  // namespace plm
