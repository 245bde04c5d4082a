// Pair data of the X path: radial functions and Y_lm with derivatives for
// the pairs of one atom, computed cooperatively by a workgroup into LDS.
#pragma once

#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "device_tables.hpp"

namespace plm {

// ---------------------------------------------------------------------------
// Pair data: radial values/derivatives and Y_lm (+ derivatives) of one pair,
// computed by one lane into LDS.
// ---------------------------------------------------------------------------
template <class T>
struct PairLayout {
    int nfn, nlmh;
    int dyw;     // elements of dY: 6 nlmh, or 0 in the values-only layout
    int stride;  // elements of T per pair slot
    // fn[nfn] fd[nfn] Y[2*nlmh] dY[6*nlmh] p[nlmh] q[nlmh] misc[8]; without
    // derivatives (deriv false: pairs_coop then writes no fd and no dY) the
    // slot has no dY
    __device__ __host__ PairLayout(int nfn_, int nlmh_, bool deriv = true)
        : nfn(nfn_), nlmh(nlmh_), dyw(deriv ? 6 * nlmh_ : 0) {
        stride = 2 * nfn + 4 * nlmh + dyw + 8;
    }
    __device__ T* fn(T* b) const { return b; }
    __device__ T* fd(T* b) const { return b + nfn; }
    __device__ Cpx<T>* Y(T* b) const { return reinterpret_cast<Cpx<T>*>(b + 2 * nfn); }
    __device__ Cpx<T>* dY(T* b) const { return reinterpret_cast<Cpx<T>*>(b + 2 * nfn + 2 * nlmh); }
    __device__ T* p(T* b) const { return b + 2 * nfn + 2 * nlmh + dyw; }
    __device__ T* q(T* b) const { return b + 2 * nfn + 3 * nlmh + dyw; }
    __device__ T* misc(T* b) const { return b + 2 * nfn + 4 * nlmh + dyw; }
    __device__ const T* misc(const T* b) const { return b + 2 * nfn + 4 * nlmh + dyw; }
};
// misc: [0..2] d, [3] r, [4] valid (1/0), [5] tp, [6] j (atom index within
// its structure)

// Pair record in device memory, written once per batch and read by every
// column block of the X rows:
//   [0..7] misc (as above), [8, 8+nfn) f_n, [8+nfn, 8+2nfn) f_n',
//   then per Y key k: Y re, im, dY/dx re, im, dY/dy re, im, dY/dz re, im.
// A block that needs keys below nkey reads only the first
// prefix(nkey) elements of a record.
struct PairRec {
    int nfn, nkey, stride;
    __device__ __host__ PairRec(int nfn_, int nkey_) : nfn(nfn_), nkey(nkey_) {
        stride = (prefix(nkey_) + 3) / 4 * 4;
    }
    __device__ __host__ int prefix(int nk) const { return 8 + 2 * nfn + 8 * nk; }
    __device__ __host__ int fn() const { return 8; }
    __device__ __host__ int fd() const { return 8 + nfn; }
    __device__ __host__ int ylm(int key) const { return 8 + 2 * nfn + 8 * key; }
};

// ---------------------------------------------------------------------------
// Cooperative pair data for a chunk of np pairs: all threads of the block
// take part (three phases separated by barriers).
//   phase 1: (pair, n) items -> misc + radial functions
//   phase 2: (pair, m) items -> Legendre column m (P and Q) with the same
//            recurrences and operation order as the reference
//   phase 3: (pair, m) items -> Y_l^{-m}, dY_l^{-m} for l = m..lmax
// dir_sign = +1 uses d = r_j - r_i + T, -1 the reversed vector.
// ---------------------------------------------------------------------------
template <class T>
__device__ void pairs_coop(const ModelDev& md, const PairLayout<T>& L, T* pairs, int np, int pc,
                           const BatchDev& bd, int t1, int a0, bool deriv, double dir_sign,
                           int lmax) {
    using C = MathConst<T>;
    const int tid = threadIdx.x, nth = blockDim.x;
    const int nfn = md.n_fn;
    // phase 1
    for (int w = tid; w < np * nfn; w += nth) {
        const int p = w / nfn, n = w % nfn;
        const int k = pc + p;
        T* slot = pairs + (size_t)p * L.stride;
        const int j = bd.nb_j[k];
        const int tp = md.tp_of[t1 * md.n_type + bd.types[j]];
        const double* dk = bd.nb_d + 3 * (size_t)k;
        const double dx = dir_sign * dk[0], dy = dir_sign * dk[1], dz = dir_sign * dk[2];
        const double dis = sqrt(dx * dx + dy * dy + dz * dz);
        if (n == 0) {
            T* m = L.misc(slot);
            m[0] = T(dx), m[1] = T(dy), m[2] = T(dz), m[3] = T(dis);
            m[4] = dis < md.cutoff ? T(1) : T(0);
            m[5] = T(tp), m[6] = T(j - a0);
        }
        if (dis >= md.cutoff || n >= md.rad_count[tp]) continue;
        const double* rp = md.rad + ((size_t)tp * md.n_fn + n) * 2;
        const T r = T(dis), rc = T(md.cutoff);
        const T kk = C::pi / rc;
        const T beta = T(rp[0]), mu = T(rp[1]);
        const T x = r - mu;
        if (deriv) {
            T sv, cv;
            dsincos(kk * r, &sv, &cv);
            const T fc = T(0.5) * (cv + T(1));
            const T fcd = T(-0.5) * kk * sv;
            const T g = dexp(-beta * x * x);
            const T gd = T(-2) * beta * x * g;
            L.fn(slot)[n] = g * fc;
            L.fd(slot)[n] = gd * fc + g * fcd;
        } else {
            const T fc = T(0.5) * (cos(kk * r) + T(1));
            L.fn(slot)[n] = dexp(-beta * x * x) * fc;
        }
    }
    if (lmax < 0) {
        __syncthreads();
        return;
    }
    __syncthreads();
    const int G = lmax + 1;
    // phase 2: Legendre column m
    for (int w = tid; w < np * G; w += nth) {
        const int p = w / G, m = w % G;
        T* slot = pairs + (size_t)p * L.stride;
        const T* ms = L.misc(slot);
        if (ms[4] == T(0)) continue;
        const T ct = ms[2] / ms[3];
        const T st = dsqrt(T(1) - ct * ct);
        T* P = L.p(slot);
        T* Q = L.q(slot);
        T pmm = C::sqrt_inv_2pi, qmm = T(0);
        if (m >= 1) {
            pmm = -st * C::sqrt3div2 * C::sqrt_inv_2pi;
            qmm = -C::sqrt3div2 * C::sqrt_inv_2pi;
            for (int l = 2; l <= m; ++l) {
                const T c1 = -dsqrt(T(1) + T(0.5) / T(l)) * st;
                pmm = c1 * pmm;
                qmm = c1 * qmm;
            }
        }
        P[lm2i(m, m)] = pmm;
        Q[lm2i(m, m)] = qmm;
        if (m + 1 <= lmax) {
            T p1, q1;
            if (m == 0) {
                p1 = ct * C::sqrt3 * C::sqrt_inv_2pi;
                q1 = T(0);
            } else {
                const T c2 = dsqrt(T(2) * (T(m + 1) - T(1)) + T(3)) * ct;
                p1 = c2 * pmm;
                q1 = c2 * qmm;
            }
            P[lm2i(m + 1, m)] = p1;
            Q[lm2i(m + 1, m)] = q1;
            T pa = pmm, pb = p1, qa = qmm, qb = q1;  // (l-2, l-1)
            for (int l = m + 2; l <= lmax; ++l) {
                const T ls = T(l * l), lm1s = T((l - 1) * (l - 1)), msq = T(m * m);
                const T alm = dsqrt((T(4) * ls - T(1)) / (ls - msq));
                const T blm = -dsqrt((lm1s - msq) / (T(4) * lm1s - T(1)));
                const T pn = alm * (ct * pb + blm * pa);
                const T qn = alm * (ct * qb + blm * qa);
                P[lm2i(l, m)] = pn;
                Q[lm2i(l, m)] = qn;
                pa = pb, pb = pn, qa = qb, qb = qn;
            }
        }
    }
    __syncthreads();
    // phase 3: Y and dY for column m
    for (int w = tid; w < np * G; w += nth) {
        const int p = w / G, m = w % G;
        T* slot = pairs + (size_t)p * L.stride;
        const T* ms = L.misc(slot);
        if (ms[4] == T(0)) continue;
        const T x = ms[0], y = ms[1], z = ms[2], r = ms[3];
        const T ct = z / r;
        const T rho = dhypot(x, y);
        T cp = T(1), sp = T(0);
        if (rho > T(0)) cp = x / rho, sp = y / rho;
        const T st = dsqrt(T(1) - ct * ct);
        const T invr = T(1) / r;
        const T* P = L.p(slot);
        const T* Q = L.q(slot);
        Cpx<T>* Y = L.Y(slot);
        Cpx<T>* dY = L.dY(slot);
        if (m == 0) {
            for (int l = 0; l <= lmax; ++l) {
                const int key = lm2i(l, 0) + l;
                Y[key] = {P[lm2i(l, 0)] * C::half_sqrt2, T(0)};
                if (deriv) {
                    const T common =
                        (l >= 1) ? Q[lm2i(l, 1)] * st * invr * dsqrt(T(0.5) * T(l) * T(l + 1)) : T(0);
                    dY[3 * key + 0] = {common * ct * cp, T(0)};
                    dY[3 * key + 1] = {common * ct * sp, T(0)};
                    dY[3 * key + 2] = {-common * st, T(0)};
                }
            }
            continue;
        }
        // cos / sin (m phi) by the Chebyshev recurrence, sign = (-1)^m
        T c1 = T(1), c2 = cp, s1 = T(0), s2 = -sp;
        const T tc = T(2) * c2;
        T c = T(1), s = T(0);
        for (int mp = 1; mp <= m; ++mp) {
            s = tc * s1 - s2;
            c = tc * c1 - c2;
            c2 = c1, c1 = c, s2 = s1, s1 = s;
        }
        const T sign = (m % 2 == 1) ? T(-1) : T(1);
        for (int l = m; l <= lmax; ++l) {
            const int key = lm2i(l, -m) + l;
            const T tmp = sign * P[lm2i(l, m)] * C::half_sqrt2;
            Y[key] = {tmp * c, -tmp * s};
            if (deriv) {
                const T cre = c * C::half_sqrt2 * invr, cim = s * C::half_sqrt2 * invr;
                T dth = T(m) * ct * Q[lm2i(l, m)];
                if (m != l) dth += dsqrt(T((l - m) * (l + m + 1))) * Q[lm2i(l, m + 1)] * st;
                const T dph = T(m) * Q[lm2i(l, m)];
                const T ure = dth * ct * cp, uim = -dph * sp;
                const T vre = dth * ct * sp, vim = dph * cp;
                const T wre = -dth * st;
                dY[3 * key + 0] = {sign * (cre * ure - cim * uim), -sign * (cre * uim + cim * ure)};
                dY[3 * key + 1] = {sign * (cre * vre - cim * vim), -sign * (cre * vim + cim * vre)};
                dY[3 * key + 2] = {sign * (cre * wre), -sign * (cim * wre)};
            }
        }
    }
    __syncthreads();
}

}  // namespace plm
