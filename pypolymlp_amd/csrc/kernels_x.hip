// HIP kernels of the X path for gfx950: neighbour lists, order parameters,
// and the energy/stress/force rows of X built per (structure, column block)
// with the block's rows resident in LDS.
//
// Reference loop nests these kernels replace:
//   NeighborFull          src/cxx/src/compute/neighbor_full.cpp:10-76
//   Local::compute_anlmtp_d (values, conjugate fill)  local.cpp:116-217
//   Features::compute_features / compute_features_deriv  polymlp_features.cpp:173-248
//   Model::model_polynomial / model_order{1,2,3}          model.cpp:159-267
//   PyModel row layout (per structure: energy, 6 stress, 3N forces)  py_model.cpp:58-106
#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "device_tables.hpp"
#include "kernels.hpp"

namespace plm {

// ---------------------------------------------------------------------------
// Neighbour list: one thread per atom i, brute force over (j, T) in the
// reference's loop order (j ascending, translations in host order). The
// distance predicate uses explicitly rounded FP64 operations so the pair set is
// bit-identical to the reference's (no FMA contraction).
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool pair_test(const double* pi, const double* pj, const double* tr,
                                          double cut2, double tol2, double* d) {
    const double dxij = __dadd_rn(pj[0], -pi[0]);
    const double dyij = __dadd_rn(pj[1], -pi[1]);
    const double dzij = __dadd_rn(pj[2], -pi[2]);
    const double dx = __dadd_rn(dxij, tr[0]);
    const double dy = __dadd_rn(dyij, tr[1]);
    const double dz = __dadd_rn(dzij, tr[2]);
    const double r2 = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
    d[0] = dx, d[1] = dy, d[2] = dz;
    return r2 < cut2 && r2 > tol2;
}

__global__ void neighbor_full_kernel(int n_atoms_total, const int* __restrict__ atom_st,
                                     const int* __restrict__ st_atom_off,
                                     const double* __restrict__ pos,
                                     const int* __restrict__ tr_off,
                                     const double* __restrict__ trans, double cut2, double tol2,
                                     int* __restrict__ count, const int* __restrict__ nb_off,
                                     int* __restrict__ nb_j, int* __restrict__ nb_t,
                                     double* __restrict__ nb_d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_atoms_total) return;
    const int s = atom_st[i];
    const int a0 = st_atom_off[s], a1 = st_atom_off[s + 1];
    const int t0 = tr_off[s], t1 = tr_off[s + 1];
    const double* pi = pos + 3 * (size_t)i;
    int c = 0;
    int out = nb_off ? nb_off[i] : 0;
    for (int j = a0; j < a1; ++j) {
        const double* pj = pos + 3 * (size_t)j;
        for (int t = t0; t < t1; ++t) {
            double d[3];
            if (pair_test(pi, pj, trans + 3 * (size_t)t, cut2, tol2, d)) {
                if (nb_off) {
                    nb_j[out] = j;
                    nb_t[out] = t - t0;
                    nb_d[3 * (size_t)out] = d[0];
                    nb_d[3 * (size_t)out + 1] = d[1];
                    nb_d[3 * (size_t)out + 2] = d[2];
                    ++out;
                }
                ++c;
            }
        }
    }
    if (!nb_off) count[i] = c;
}

// Half list (neighbor_half.cpp:10-100): j < i, plus self images in the upper
// half space of the translation.
__global__ void neighbor_half_kernel(int n_atoms_total, const int* __restrict__ atom_st,
                                     const int* __restrict__ st_atom_off,
                                     const double* __restrict__ pos,
                                     const int* __restrict__ tr_off,
                                     const double* __restrict__ trans, double cut2, double tol2,
                                     double tol, int* __restrict__ count,
                                     const int* __restrict__ nb_off, int* __restrict__ nb_j,
                                     int* __restrict__ nb_t, double* __restrict__ nb_d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_atoms_total) return;
    const int s = atom_st[i];
    const int a0 = st_atom_off[s];
    const int t0 = tr_off[s], t1 = tr_off[s + 1];
    const double* pi = pos + 3 * (size_t)i;
    int c = 0;
    int out = nb_off ? nb_off[i] : 0;
    auto emit = [&](int j, int t, const double* d) {
        if (nb_off) {
            nb_j[out] = j;
            nb_t[out] = t - t0;
            nb_d[3 * (size_t)out] = d[0];
            nb_d[3 * (size_t)out + 1] = d[1];
            nb_d[3 * (size_t)out + 2] = d[2];
            ++out;
        }
        ++c;
    };
    for (int j = a0; j < i; ++j) {
        const double* pj = pos + 3 * (size_t)j;
        for (int t = t0; t < t1; ++t) {
            double d[3];
            if (pair_test(pi, pj, trans + 3 * (size_t)t, cut2, tol2, d)) emit(j, t, d);
        }
    }
    for (int t = t0; t < t1; ++t) {
        const double* tr = trans + 3 * (size_t)t;
        const double dx = tr[0], dy = tr[1], dz = tr[2];
        const double r2 =
            __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
        if (r2 < cut2 && r2 > tol2) {
            bool keep;
            if (dz >= tol) keep = true;
            else if (fabs(dz) < tol && dy >= tol) keep = true;
            else if (fabs(dz) < tol && fabs(dy) < tol && dx >= tol) keep = true;
            else keep = false;
            if (keep) {
                double d[3] = {dx, dy, dz};
                emit(i, t, d);
            }
        }
    }
    if (!nb_off) count[i] = c;
}

// ---------------------------------------------------------------------------
// Pair data: radial values/derivatives and Y_lm (+ derivatives) of one pair,
// computed by one lane into LDS.
// ---------------------------------------------------------------------------
template <class T>
struct PairLayout {
    int nfn, nlmh;
    int stride;  // elements of T per pair slot
    __device__ __host__ PairLayout(int nfn_, int nlmh_) : nfn(nfn_), nlmh(nlmh_) {
        // fn[nfn] fd[nfn] Y[2*nlmh] dY[6*nlmh] p[nlmh] q[nlmh] misc[8]
        stride = 2 * nfn + 10 * nlmh + 8;
    }
    __device__ T* fn(T* b) const { return b; }
    __device__ T* fd(T* b) const { return b + nfn; }
    __device__ Cpx<T>* Y(T* b) const { return reinterpret_cast<Cpx<T>*>(b + 2 * nfn); }
    __device__ Cpx<T>* dY(T* b) const { return reinterpret_cast<Cpx<T>*>(b + 2 * nfn + 2 * nlmh); }
    __device__ T* p(T* b) const { return b + 2 * nfn + 8 * nlmh; }
    __device__ T* q(T* b) const { return b + 2 * nfn + 9 * nlmh; }
    __device__ T* misc(T* b) const { return b + 2 * nfn + 10 * nlmh; }
    __device__ const T* misc(const T* b) const { return b + 2 * nfn + 10 * nlmh; }
};

// misc: [0..2] d, [3] r, [4] valid (1/0), [5] tp, [6] j (local atom index)
template <class T>
__device__ void compute_pair(const ModelDev& md, const PairLayout<T>& L, T* slot, const double* d,
                             int tp, int jloc, bool deriv, int lmax) {
    T* m = L.misc(slot);
    const double dx = d[0], dy = d[1], dz = d[2];
    const double dis = sqrt(dx * dx + dy * dy + dz * dz);
    m[0] = T(dx), m[1] = T(dy), m[2] = T(dz), m[3] = T(dis);
    m[5] = T(tp), m[6] = T(jloc);
    if (dis >= md.cutoff) {
        m[4] = T(0);
        return;
    }
    m[4] = T(1);
    const int cnt = md.rad_count[tp];
    const double* rad = md.rad + (size_t)tp * md.n_fn * 2;
    if (deriv) radial<T>(T(dis), T(md.cutoff), rad, cnt, L.fn(slot), L.fd(slot));
    else radial_value<T>(T(dis), T(md.cutoff), rad, cnt, L.fn(slot));
    if (lmax >= 0)
        ylm<T>(lmax, T(dx), T(dy), T(dz), T(dis), L.Y(slot), deriv ? L.dY(slot) : nullptr,
               L.p(slot), L.q(slot));
}

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

// ---------------------------------------------------------------------------
// Order parameters a_{nlm,tp}(i) (noconj accumulation + conjugate fill).
// One workgroup per atom. skip_tol follows local.cpp:164 (f_n < 1e-20 skipped).
// ---------------------------------------------------------------------------
constexpr int kAnlmPairs = 16;

template <class T>
__global__ void __launch_bounds__(256) anlm_kernel(ModelDev md, BatchDev bd,
                                                   Cpx<T>* __restrict__ a_full, int a_stride,
                                                   int atom0, int n_atoms, int negate) {
    extern __shared__ unsigned char smem_raw[];
    const int ia = blockIdx.x;
    if (ia >= n_atoms) return;
    const int i = atom0 + ia;
    const int t1 = bd.types[i];
    const TypeDev& TT = md.t[t1];
    const PairLayout<T> L(md.n_fn, md.n_lmh);
    const int NPS = kAnlmPairs;  // pairs per chunk, one lane each
    T* pairs = reinterpret_cast<T*>(smem_raw);
    Cpx<T>* acc = reinterpret_cast<Cpx<T>*>(pairs + (size_t)NPS * L.stride);
    for (int h = threadIdx.x; h < TT.n_noconj; h += blockDim.x) acc[h] = {T(0), T(0)};
    const int p0 = bd.nb_off[i], p1 = bd.nb_off[i + 1];
    const int lmax = md.max_l;
    for (int c0 = p0; c0 < p1; c0 += NPS) {
        const int np = min(NPS, p1 - c0);
        __syncthreads();
        // X path: d = r_j - r_i + T (local.cpp:149); evaluation: the reversed
        // vector r_i - r_j - T (polymlp_eval.cpp:376-379)
        pairs_coop<T>(md, L, pairs, np, c0, bd, t1, 0, false, negate ? -1.0 : 1.0, lmax);
        for (int p = 0; p < np; ++p) {
            T* slot = pairs + (size_t)p * L.stride;
            const T* m = L.misc(slot);
            if (m[4] == T(0)) continue;
            const int tp = int(m[5]);
            const T* fn = L.fn(slot);
            const Cpx<T>* Y = L.Y(slot);
            const int h0 = TT.nc_tp_off[tp], h1 = TT.nc_tp_off[tp + 1];
            for (int h = h0 + threadIdx.x; h < h1; h += blockDim.x) {
                const T f = fn[TT.nc_tp_nid[h]];
                if (f < T(1e-20)) continue;
                const int key = TT.nc_tp_ylm[h];
                const int hd = TT.nc_tp_head[h];
                if (key < 0) {
                    acc[hd].re += f;
                } else {
                    acc[hd].re += f * Y[key].re;
                    acc[hd].im += f * Y[key].im;
                }
            }
        }
    }
    __syncthreads();
    Cpx<T>* out = a_full + (size_t)ia * a_stride;
    for (int k = threadIdx.x; k < TT.n_nlm; k += blockDim.x) {
        const Cpx<T> v = acc[TT.nlm_src[k]];
        if (TT.nlm_conj[k]) {
            const T cc = T(TT.nlm_sign[k]);
            out[k] = {cc * v.re, -cc * v.im};
        } else {
            out[k] = v;
        }
    }
}

// Product of up to `nk` order parameters (keys -1 terminate).
template <class T>
__device__ __forceinline__ Cpx<T> prod_keys(const Cpx<T>* a, const int* keys, int nk) {
    Cpx<T> v = {T(1), T(0)};
    for (int k = 0; k < nk; ++k) {
        const int id = keys[k];
        if (id < 0) break;
        v = cmul(v, a[id]);
    }
    return v;
}

// ---------------------------------------------------------------------------
// X rows of one (structure, column block): energy row, 6 stress rows and 3N
// force rows of the block's columns, accumulated in LDS and written once.
//
// Column c = prod_k x_{b_k}; for atom i of type t and neighbour pair p=(i->j):
//   dx_f/dd = Re sum_h W'_i[f,h] (f'_n Y_lm d/r + f_n dY_lm)   (h = (n,l,m,tp_p))
//   W'_i[f,h] = sum over invariant terms of coeff * prod(other factors)
// Force rows: +q to atom i, -q to atom j; stress rows: -q (x) d.
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(256) xblock_kernel(ModelDev md, ColumnsDev cd, BatchDev bd,
                                                     BlockWz bw, XBlockParams P,
                                                     const Cpx<T>* __restrict__ a_full,
                                                     int a_stride, int atom_base,
                                                     T* __restrict__ X,
                                                     const double* __restrict__ roww,
                                                     double* __restrict__ xe_sum,
                                                     double* __restrict__ xe_sq) {
    extern __shared__ unsigned char smem_raw[];
    const int blk = blockIdx.x % cd.n_blocks;
    const int s = P.st0 + blockIdx.x / cd.n_blocks;
    const int c0 = cd.blk_c0[blk];
    const int ncol = cd.blk_c0[blk + 1] - c0;
    const int a0 = bd.st_atom_off[s], a1 = bd.st_atom_off[s + 1];
    const int nat = a1 - a0;
    const int rows = 7 + 3 * nat;
    const int tid = threadIdx.x;
    const int ntp = md.n_tp;
    const PairLayout<T> L(md.n_fn, md.n_lmh);

    // LDS carve-up
    T* tile = reinterpret_cast<T*>(smem_raw);                       // rows_max * CB
    Cpx<T>* Wv = reinterpret_cast<Cpx<T>*>(tile + (size_t)P.rows_max * P.CB);  // wz_max
    T* xneed = reinterpret_cast<T*>(Wv + P.wz_max);                 // need_max
    T* coef = xneed + P.need_max;                                   // CB * 3
    T* qbuf = coef + 3 * P.CB;                                      // NP * need_max * 3
    T* pairs = qbuf + (size_t)P.NP * P.need_max * 3;                // NP * stride

    for (int e = tid; e < rows * P.CB; e += blockDim.x) tile[e] = T(0);

    // thread -> (column, pair group)
    const int col = tid % P.CB;
    const int pg = tid / P.CB;
    const int npg = blockDim.x / P.CB;
    T sacc[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};

    for (int ia = 0; ia < nat; ++ia) {
        const int i = a0 + ia;
        const int t1 = bd.types[i];
        const TypeDev& TT = md.t[t1];
        const Cpx<T>* a = a_full + (size_t)(i - atom_base) * a_stride;
        const int nf0 = cd.need_off[t1 * (cd.n_blocks + 1) + blk];
        const int nneed = cd.need_off[t1 * (cd.n_blocks + 1) + blk + 1] - nf0;
        const int* need = cd.need_feat + nf0;
        const int z0 = bw.off[t1 * (cd.n_blocks + 1) + blk];
        const int nz = bw.off[t1 * (cd.n_blocks + 1) + blk + 1] - z0;
        const int* needwz = bw.need_wz_off + nf0;
        __syncthreads();
        // feature values of the needed features
        for (int k = tid; k < nneed; k += blockDim.x) {
            const int f = need[k];
            T v = T(0);
            for (int e = TT.fv_off[f]; e < TT.fv_off[f + 1]; ++e) {
                const Cpx<T> pr = prod_keys<T>(a, TT.fv_keys + 4 * (size_t)e, 4);
                v += T(TT.fv_coeff[e]) * pr.re;
            }
            xneed[k] = v;
        }
        // W' values of the needed features' nonzeros
        for (int z = tid; z < nz; z += blockDim.x) {
            const int g = bw.nz[z0 + z];
            Cpx<T> w = {T(0), T(0)};
            for (int e = TT.wz_term_off[g]; e < TT.wz_term_off[g + 1]; ++e) {
                Cpx<T> pr = prod_keys<T>(a, TT.wt_keys + 3 * (size_t)e, 3);
                if (TT.wt_conj[e]) pr.im = -pr.im;
                const T c = T(TT.wt_coeff[e]);
                w.re += c * pr.re;
                w.im += c * pr.im;
            }
            Wv[z] = w;
        }
        __syncthreads();
        // per-column polynomial coefficients and energy row
        if (tid < ncol) {
            const int c = c0 + tid;
            const int nfac = cd.nfac[c];
            const int* cn = cd.col_need + ((size_t)t1 * cd.n_cols + c) * 3;
            T x0 = T(0), x1 = T(0), x2 = T(0);
            bool valid = cn[0] >= 0;
            if (valid) {
                x0 = xneed[cn[0]];
                if (nfac > 1) x1 = xneed[cn[1]];
                if (nfac > 2) x2 = xneed[cn[2]];
            }
            T e = T(0), k0 = T(0), k1 = T(0), k2 = T(0);
            if (valid) {
                if (nfac == 1) e = x0, k0 = T(1);
                else if (nfac == 2) e = x0 * x1, k0 = x1, k1 = x0;
                else e = x0 * x1 * x2, k0 = x1 * x2, k1 = x0 * x2, k2 = x0 * x1;
            }
            tile[tid] += e;
            coef[3 * tid] = k0, coef[3 * tid + 1] = k1, coef[3 * tid + 2] = k2;
        }
        T self[3] = {T(0), T(0), T(0)};
        const int p0 = bd.nb_off[i], p1 = bd.nb_off[i + 1];
        for (int pc = p0; pc < p1; pc += P.NP) {
            const int np = min(P.NP, p1 - pc);
            __syncthreads();
            pairs_coop<T>(md, L, pairs, np, pc, bd, t1, a0, true, 1.0, P.lmax);
            // q[p][k][alpha] for needed features
            for (int w = tid; w < np * nneed; w += blockDim.x) {
                const int p = w / nneed, k = w % nneed;
                T* slot = pairs + (size_t)p * L.stride;
                const T* m = L.misc(slot);
                T q0 = T(0), q1 = T(0), q2 = T(0);
                if (m[4] != T(0)) {
                    const int tp = int(m[5]);
                    const int f = need[k];
                    const int g0 = TT.wz_off[f * ntp], ga = TT.wz_off[f * ntp + tp],
                              gb = TT.wz_off[f * ntp + tp + 1];
                    const int zb = needwz[k] + (ga - g0);
                    const T* fn = L.fn(slot);
                    const T* fd = L.fd(slot);
                    const Cpx<T>* Y = L.Y(slot);
                    const Cpx<T>* dY = L.dY(slot);
                    const T ir = T(1) / m[3];
                    const T ux = m[0] * ir, uy = m[1] * ir, uz = m[2] * ir;
                    for (int g = ga; g < gb; ++g) {
                        const int n = TT.wz_nid[g];
                        const T fv = fn[n];
                        if (fv < T(1e-20)) continue;
                        const T fdv = fd[n];
                        const Cpx<T> W = Wv[zb + (g - ga)];
                        const int key = TT.wz_ylm[g];
                        if (key < 0) {
                            const T r0 = fdv * W.re;
                            q0 += r0 * ux, q1 += r0 * uy, q2 += r0 * uz;
                        } else {
                            const T ry = re_mul(W, Y[key]) * fdv;
                            q0 += ry * ux + fv * re_mul(W, dY[3 * key]);
                            q1 += ry * uy + fv * re_mul(W, dY[3 * key + 1]);
                            q2 += ry * uz + fv * re_mul(W, dY[3 * key + 2]);
                        }
                    }
                }
                T* qq = qbuf + ((size_t)p * P.need_max + k) * 3;
                qq[0] = q0, qq[1] = q1, qq[2] = q2;
            }
            __syncthreads();
            if (col < ncol) {
                const int c = c0 + col;
                const int* cn = cd.col_need + ((size_t)t1 * cd.n_cols + c) * 3;
                if (cn[0] >= 0) {
                    const int nfac = cd.nfac[c];
                    const T k0 = coef[3 * col], k1 = coef[3 * col + 1], k2 = coef[3 * col + 2];
                    for (int p = pg; p < np; p += npg) {
                        const T* slot = pairs + (size_t)p * L.stride;
                        const T* m = L.misc(slot);
                        if (m[4] == T(0)) continue;
                        const T* q = qbuf + (size_t)p * P.need_max * 3;
                        T v0 = k0 * q[3 * cn[0]], v1 = k0 * q[3 * cn[0] + 1],
                          v2 = k0 * q[3 * cn[0] + 2];
                        if (nfac > 1) {
                            v0 += k1 * q[3 * cn[1]], v1 += k1 * q[3 * cn[1] + 1],
                                v2 += k1 * q[3 * cn[1] + 2];
                        }
                        if (nfac > 2) {
                            v0 += k2 * q[3 * cn[2]], v1 += k2 * q[3 * cn[2] + 1],
                                v2 += k2 * q[3 * cn[2] + 2];
                        }
                        self[0] += v0, self[1] += v1, self[2] += v2;
                        const int jl = int(m[6]);
                        T* rj = tile + (size_t)(7 + 3 * jl) * P.CB + col;
                        atomicAdd(rj, -v0);
                        atomicAdd(rj + P.CB, -v1);
                        atomicAdd(rj + 2 * P.CB, -v2);
                        const T dx = m[0], dy = m[1], dz = m[2];
                        sacc[0] -= v0 * dx, sacc[1] -= v1 * dy, sacc[2] -= v2 * dz;
                        sacc[3] -= v0 * dy, sacc[4] -= v1 * dz, sacc[5] -= v2 * dx;
                    }
                }
            }
        }
        if (col < ncol) {
            T* ri = tile + (size_t)(7 + 3 * ia) * P.CB + col;
            atomicAdd(ri, self[0]);
            atomicAdd(ri + P.CB, self[1]);
            atomicAdd(ri + 2 * P.CB, self[2]);
        }
    }
    if (col < ncol)
        for (int k = 0; k < 6; ++k) atomicAdd(tile + (size_t)(1 + k) * P.CB + col, sacc[k]);
    __syncthreads();
    // write out (weights applied per row); raw energy row feeds the scale sums
    const int rbase = bd.st_row_off[s] - P.row_base;
    for (int e = tid; e < rows * P.CB; e += blockDim.x) {
        const int r = e / P.CB, c = e % P.CB;
        if (c >= ncol) continue;
        const T v = tile[e];
        const double w = roww ? roww[rbase + r] : 1.0;
        X[(size_t)(rbase + r) * P.ld + c0 + c] = T(w * double(v));
        if (r == 0 && xe_sum) {
            atomicAdd(xe_sum + c0 + c, double(v));
            atomicAdd(xe_sq + c0 + c, double(v) * double(v));
        }
    }
}

// explicit instantiation via launchers
void launch_neighbor_full(hipStream_t st, int n, const int* atom_st, const int* st_atom_off,
                          const double* pos, const int* tr_off, const double* trans, double cut2,
                          double tol2, int* count, const int* nb_off, int* nb_j, int* nb_t,
                          double* nb_d) {
    const int B = 128;
    neighbor_full_kernel<<<(n + B - 1) / B, B, 0, st>>>(n, atom_st, st_atom_off, pos, tr_off,
                                                        trans, cut2, tol2, count, nb_off, nb_j,
                                                        nb_t, nb_d);
}

void launch_neighbor_half(hipStream_t st, int n, const int* atom_st, const int* st_atom_off,
                          const double* pos, const int* tr_off, const double* trans, double cut2,
                          double tol2, double tol, int* count, const int* nb_off, int* nb_j,
                          int* nb_t, double* nb_d) {
    const int B = 128;
    neighbor_half_kernel<<<(n + B - 1) / B, B, 0, st>>>(n, atom_st, st_atom_off, pos, tr_off,
                                                        trans, cut2, tol2, tol, count, nb_off,
                                                        nb_j, nb_t, nb_d);
}

size_t anlm_smem(const ModelDev& md, int elem, int nthreads, int max_noconj) {
    const PairLayout<double> L(md.n_fn, md.n_lmh);
    return (size_t)nthreads * L.stride * elem + (size_t)max_noconj * 2 * elem;
}

template <class T>
void launch_anlm_t(hipStream_t st, const ModelDev& md, const BatchDev& bd, void* a_full,
                   int a_stride, int atom0, int n_atoms, int max_noconj, int negate) {
    const int B = 64;
    const size_t sm = anlm_smem(md, sizeof(T), kAnlmPairs, max_noconj);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&anlm_kernel<T>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    anlm_kernel<T><<<n_atoms, B, sm, st>>>(md, bd, reinterpret_cast<Cpx<T>*>(a_full), a_stride,
                                           atom0, n_atoms, negate);
}

void launch_anlm(hipStream_t st, bool fp64, const ModelDev& md, const BatchDev& bd, void* a_full,
                 int a_stride, int atom0, int n_atoms, int max_noconj, int negate) {
    if (fp64) launch_anlm_t<double>(st, md, bd, a_full, a_stride, atom0, n_atoms, max_noconj, negate);
    else launch_anlm_t<float>(st, md, bd, a_full, a_stride, atom0, n_atoms, max_noconj, negate);
}

size_t xblock_smem(const ModelDev& md, const XBlockParams& P, int elem) {
    const PairLayout<double> L(md.n_fn, md.n_lmh);
    size_t n = (size_t)P.rows_max * P.CB + 2 * (size_t)P.wz_max + P.need_max + 3 * (size_t)P.CB +
               (size_t)P.NP * P.need_max * 3 + (size_t)P.NP * L.stride;
    return n * elem;
}

template <class T>
void launch_xblock_t(hipStream_t st, const ModelDev& md, const ColumnsDev& cd, const BatchDev& bd,
                     const BlockWz& bw, const XBlockParams& P, int n_st, const void* a_full,
                     int a_stride, int atom_base, void* X, const double* roww, double* xe_sum,
                     double* xe_sq) {
    const size_t sm = xblock_smem(md, P, sizeof(T));
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&xblock_kernel<T>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    xblock_kernel<T><<<n_st * cd.n_blocks, 256, sm, st>>>(
        md, cd, bd, bw, P, reinterpret_cast<const Cpx<T>*>(a_full), a_stride, atom_base,
        reinterpret_cast<T*>(X), roww, xe_sum, xe_sq);
}

void launch_xblock(hipStream_t st, bool fp64, const ModelDev& md, const ColumnsDev& cd,
                   const BatchDev& bd, const BlockWz& bw, const XBlockParams& P, int n_st,
                   const void* a_full, int a_stride, int atom_base, void* X, const double* roww,
                   double* xe_sum, double* xe_sq) {
    if (fp64)
        launch_xblock_t<double>(st, md, cd, bd, bw, P, n_st, a_full, a_stride, atom_base, X, roww,
                                xe_sum, xe_sq);
    else
        launch_xblock_t<float>(st, md, cd, bd, bw, P, n_st, a_full, a_stride, atom_base, X, roww,
                               xe_sum, xe_sq);
}

}  // namespace plm
