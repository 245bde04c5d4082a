// Energy, stress and force rows of X for gfx950.
//
// pairdata_kernel evaluates, once per batch, the radial functions and Y_lm
// with their derivatives for every neighbour pair (PairRec records).
//
// xrows_kernel builds the rows of one column block of one structure with all
// of the block's rows resident in LDS. For each atom i of the structure the
// workgroup first evaluates the invariants and derivative weights the block
// needs:
//   x_f(i)        feature values (polynomial columns only)
//   W'_i[f, h]    = d x_f / d a_h   (h = non-conjugate order parameter)
// Then each wave takes its own pairs (i -> j) of atom i, stages their records
// in a private LDS slot, and every lane, which owns one column, adds
//   q = dx_c/dd = Re sum_h coef * W'_i[f, h] (f'_n Y_lm d/r + f_n dY_lm)
// to the rows of atom j (LDS atomics; consecutive lanes hit consecutive
// words), to a register for atom i and to six register stress sums. The pair
// loop itself has no workgroup barrier.
//
// Reference loop nests:
//   Local::compute_anlmtp_d        src/cxx/src/compute/local.cpp:116-217
//   Features::compute_features_deriv  src/cxx/src/polymlp/polymlp_features.cpp:208-248
//   Model::model_polynomial, model_order{1,2,3}  src/cxx/src/compute/model.cpp:159-267
//   PyModel row layout (energy, 6 stress, 3N forces)  src/cxx/src/compute/py_model.cpp:58-106
#include <hip/hip_runtime.h>

#include <stdexcept>

#include "device_math.hpp"
#include "device_tables.hpp"
#include "kernels.hpp"
#include "pair_data.hpp"

namespace plm {

constexpr int kPdPairs = 4;    // pairs per LDS chunk in pairdata_kernel: 8 KB of LDS in FP32, about 20 waves per CU (16: 5 waves, 2.1x slower)
constexpr int kMaxPairTypes = 64;  // type pairs the grouped record order supports
// staged record elements per lane (64 lanes): FP32 groups are one record
// (<= 768 values), FP64 groups two
template <class T> struct StageRegs { static constexpr int n = sizeof(T) == 8 ? 24 : 12; };

// ---------------------------------------------------------------------------
// Pair records (X path direction d = r_j - r_i + T, local.cpp:149).
// One workgroup per atom.
// ---------------------------------------------------------------------------
// One wave per atom (64 threads). With tpo, the records are stored grouped by
// type pair: lane tp keeps the next free index of type pair tp, and each
// chunk's destinations come from ballots (stable, deterministic).
template <class T>
__global__ void __launch_bounds__(64) pairdata_kernel(ModelDev md, BatchDev bd, int atom0,
                                                      int n_atoms, long pair_base,
                                                      T* __restrict__ pd, PairRec R, int lmax,
                                                      int* __restrict__ tpo) {
    extern __shared__ unsigned char smem_raw[];
    __shared__ int cnt[kMaxPairTypes];
    __shared__ int dest[kPdPairs];
    const int ia = blockIdx.x;
    if (ia >= n_atoms) return;
    const int i = atom0 + ia;
    const int t1 = bd.types[i];
    const int a0 = bd.st_atom_off[bd.atom_st[i]];
    const PairLayout<T> L(md.n_fn, md.n_lmh);
    T* pairs = reinterpret_cast<T*>(smem_raw);
    const int p0 = bd.nb_off[i], p1 = bd.nb_off[i + 1];
    const int nfn = md.n_fn;
    const int pre = R.prefix(R.nkey);
    const int tid = threadIdx.x, ntp = md.n_tp;
    int run = 0;  // lane tp: next free pair index of type pair tp
    if (tpo) {
        if (tid < ntp) cnt[tid] = 0;
        __syncthreads();
        for (int k = p0 + tid; k < p1; k += 64)
            atomicAdd(&cnt[md.tp_of[t1 * md.n_type + bd.types[bd.nb_j[k]]]], 1);
        __syncthreads();
        if (tid < ntp) {
            run = p0;
            for (int t = 0; t < tid; ++t) run += cnt[t];
            tpo[(size_t)ia * (ntp + 1) + tid] = run;
        }
        if (tid == 0) tpo[(size_t)ia * (ntp + 1) + ntp] = p1;
    }
    for (int c0 = p0; c0 < p1; c0 += kPdPairs) {
        const int np = min(kPdPairs, p1 - c0);
        __syncthreads();
        pairs_coop<T>(md, L, pairs, np, c0, bd, t1, a0, true, 1.0, lmax);
        if (tpo) {
            const int myt = tid < np ? int(L.misc(pairs + (size_t)tid * L.stride)[5]) : -1;
            int d = 0;
            for (int t = 0; t < ntp; ++t) {
                const unsigned long long m = __ballot(myt == t);
                const int base = __shfl(run, t);
                if (myt == t) d = base + __popcll(m & ((1ull << tid) - 1ull));
                if (tid == t) run += __popcll(m);
            }
            if (tid < np) dest[tid] = d;
            __syncthreads();
        }
        for (int e = threadIdx.x; e < np * pre; e += blockDim.x) {
            const int p = e / pre, o = e % pre;
            T* slot = pairs + (size_t)p * L.stride;
            T v = T(0);
            if (o < 7) {
                v = L.misc(slot)[o];
            } else if (o >= 8 && o < 8 + nfn) {
                v = L.fn(slot)[o - 8];
            } else if (o >= 8 + nfn && o < 8 + 2 * nfn) {
                v = L.fd(slot)[o - 8 - nfn];
            } else if (o >= 8 + 2 * nfn) {
                const int q = o - 8 - 2 * nfn;
                const int key = q >> 3, c = q & 7;
                if (c < 2) {
                    const Cpx<T> y = L.Y(slot)[key];
                    v = c == 0 ? y.re : y.im;
                } else {
                    const Cpx<T> y = L.dY(slot)[3 * key + ((c - 2) >> 1)];
                    v = (c & 1) ? y.im : y.re;
                }
            }
            // pairs beyond the cutoff keep only misc (valid = 0)
            if (o >= 8 && L.misc(slot)[4] == T(0)) v = T(0);
            const long dp = tpo ? dest[p] : c0 + p;
            pd[(size_t)(dp - pair_base) * R.stride + o] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// Order parameters of the X path from the pair records (records grouped by
// type pair, tpo): a_h = sum over the pairs of head h's type pair of
// f_n(r) Y_lm(d), pairs with f_n < 1e-20 skipped (local.cpp:164), in the
// neighbour order within the type pair; then the atom's full component list
// with the conjugates (Local::compute_anlmtp, local.cpp:62-110). One
// workgroup per atom; chunks of pairs staged in LDS.
// ---------------------------------------------------------------------------
constexpr int kArPairs = 32;  // pairs per LDS chunk in anlm_rec_kernel

template <class T>
__global__ void __launch_bounds__(256) anlm_rec_kernel(ModelDev md, BatchDev bd, int atom0,
                                                       int n_atoms, long pair_base,
                                                       const T* __restrict__ pd, PairRec R,
                                                       const int* __restrict__ tpo,
                                                       Cpx<T>* __restrict__ a_full, int a_stride) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    const int ia = blockIdx.x;
    if (ia >= n_atoms) return;
    const int i = atom0 + ia;
    const TypeDev& TT = md.t[bd.types[i]];
    const int nfn = md.n_fn, nlmh = md.n_lmh, ntp = md.n_tp;
    const int W = nfn + 2 * nlmh;  // staged values per pair: f_n, then (Re, Im) Y per key
    T* fL = reinterpret_cast<T*>(smem_raw);
    Cpx<T>* acc = reinterpret_cast<Cpx<T>*>(fL + (size_t)kArPairs * W);
    const int tid = threadIdx.x;
    for (int h = tid; h < TT.n_noconj; h += blockDim.x) acc[h] = {T(0), T(0)};
    const int* to = tpo + (size_t)ia * (ntp + 1);
    for (int tp = 0; tp < ntp; ++tp) {
        const int h0 = TT.nc_tp_off[tp], h1 = TT.nc_tp_off[tp + 1];
        if (h0 == h1) continue;
        for (int c0 = to[tp]; c0 < to[tp + 1]; c0 += kArPairs) {
            const int np = min(kArPairs, to[tp + 1] - c0);
            __syncthreads();
            // eight loads in flight per thread
            for (int e = tid; e < np * W; e += 8 * blockDim.x) {
                T v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int ee = e + u * blockDim.x;
                    if (ee >= np * W) break;
                    const int p = ee / W, o = ee % W;
                    const T* rec = pd + (size_t)(c0 + p - pair_base) * R.stride;
                    v[u] = o < nfn ? rec[R.fn() + o] : rec[R.ylm((o - nfn) >> 1) + ((o - nfn) & 1)];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int ee = e + u * blockDim.x;
                    if (ee < np * W) fL[ee] = v[u];
                }
            }
            __syncthreads();
            for (int h = h0 + tid; h < h1; h += blockDim.x) {
                const int nid = TT.nc_tp_nid[h], key = TT.nc_tp_ylm[h];
                Cpx<T> v = acc[TT.nc_tp_head[h]];
                for (int p = 0; p < np; ++p) {
                    const T* q = fL + (size_t)p * W;
                    const T f = q[nid];
                    if (f < T(1e-20)) continue;
                    if (key < 0) {
                        v.re += f;
                    } else {
                        v.re += f * q[nfn + 2 * key];
                        v.im += f * q[nfn + 2 * key + 1];
                    }
                }
                acc[TT.nc_tp_head[h]] = v;
            }
        }
    }
    __syncthreads();
    Cpx<T>* out = a_full + (size_t)ia * a_stride;
    for (int k = tid; k < TT.n_nlm; k += blockDim.x) {
        const Cpx<T> v = acc[TT.nlm_src[k]];
        if (TT.nlm_conj[k]) {
            const T cc = T(TT.nlm_sign[k]);
            out[k] = {cc * v.re, -cc * v.im};
        } else {
            out[k] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// Derivative weights W'_i[g] of every nonzero g of the atom's type and the
// feature values x_f(i), once per atom (one workgroup per atom, the order
// parameters staged in LDS).
//   W'[f, h] = sum over invariant terms c * prod(other factors), conjugate
//   heads folded in (polymlp_features.cpp:208-248, prod_anlmtp_deriv);
//   x_f = sum c * Re prod, dropping |Re prod| <= 1e-20 where the reference
//   does (polymlp_features.cpp:184-193).
// ---------------------------------------------------------------------------
template <class T> __device__ __forceinline__ const T* wq_coeff(const TypeDev& TT);
template <> __device__ __forceinline__ const double* wq_coeff<double>(const TypeDev& TT) { return TT.wq_cc; }
template <> __device__ __forceinline__ const float* wq_coeff<float>(const TypeDev& TT) { return TT.wq_ccf; }
template <class T> __device__ __forceinline__ const T* fq_coeff(const TypeDev& TT);
template <> __device__ __forceinline__ const double* fq_coeff<double>(const TypeDev& TT) { return TT.fq_c; }
template <> __device__ __forceinline__ const float* fq_coeff<float>(const TypeDev& TT) { return TT.fq_cf; }

// Products of NK order-parameter components (keys -1 padded). a[one] holds
// 1 + 0i, so padded keys multiply by one and the loop has no branch; the
// factors are taken in key order, as the reference multiplies them.
template <class T, int NK>
__device__ __forceinline__ Cpx<T> prod_keys(const Cpx<T>* a, int4 k, int one) {
    Cpx<T> v = a[k.x < 0 ? one : k.x];
    if (NK > 1) v = cmul(v, a[k.y < 0 ? one : k.y]);
    if (NK > 2) v = cmul(v, a[k.z < 0 ? one : k.z]);
    if (NK > 3) v = cmul(v, a[k.w < 0 ? one : k.w]);
    return v;
}

// NF: factors per W' term, NV: per value term (maxima over the model's types)
template <class T, int NF, int NV>
__global__ void __launch_bounds__(256) wprime_kernel(ModelDev md, BatchDev bd,
                                                     const Cpx<T>* __restrict__ a_full,
                                                     int a_stride, int atom0, int n_atoms,
                                                     Cpx<T>* __restrict__ w, int w_stride,
                                                     T* __restrict__ x, int x_stride) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    Cpx<T>* a = reinterpret_cast<Cpx<T>*>(smem_raw);
    const int ia = blockIdx.x;
    if (ia >= n_atoms) return;
    const int i = atom0 + ia;
    const int ti = bd.types[i];
    const TypeDev& TT = md.t[ti];
    const Cpx<T>* ag = a_full + (size_t)ia * a_stride;
    const int one = TT.n_nlm;
    stage_cpx(a, ag, TT.n_nlm);
    if (threadIdx.x == 0) a[one] = {T(1), T(0)};
    __syncthreads();
    Cpx<T>* wo = w + (size_t)ia * w_stride;
    const int lane = threadIdx.x & 63, nwave = blockDim.x >> 6;
    // W': one block of 64 nonzeros per wave at a time, lane l takes term row
    // j of its nonzero (rows are contiguous across the lanes)
    {
        const T* cc = wq_coeff<T>(TT);
        const int4* k4 = reinterpret_cast<const int4*>(TT.wq_keys);
        for (int bk = threadIdx.x >> 6; bk < TT.wq_nblk; bk += nwave) {
            const int g = TT.wq_item[bk * 64 + lane];
            T vr = T(0), vi = T(0);
            // the next row's key record and coefficients load while this
            // row's order parameters are gathered
            const int r1 = TT.wq_boff[bk + 1];
            int r = TT.wq_boff[bk] + lane;
            int4 kn = r < r1 ? k4[r] : int4{-1, -1, -1, -1};
            T cn0 = r < r1 ? cc[2 * r] : T(0), cn1 = r < r1 ? cc[2 * r + 1] : T(0);
            for (; r < r1; r += 64) {
                const int4 kk = kn;
                const T c0 = cn0, c1 = cn1;
                if (r + 64 < r1) kn = k4[r + 64], cn0 = cc[2 * (r + 64)], cn1 = cc[2 * (r + 64) + 1];
                const Cpx<T> pr = prod_keys<T, NF>(a, kk, one);
                vr += c0 * pr.re;
                vi += c1 * pr.im;
            }
            if (g >= 0) wo[g] = Cpx<T>{vr, vi};
        }
    }
    // feature values, same layout; products with |Re| <= 1e-20 dropped where
    // the reference drops them
    T* xo = x + (size_t)ia * x_stride;
    {
        const T* c = fq_coeff<T>(TT);
        const int4* k4 = reinterpret_cast<const int4*>(TT.fq_keys);
        for (int bk = threadIdx.x >> 6; bk < TT.fq_nblk; bk += nwave) {
            const int f = TT.fq_item[bk * 64 + lane];
            T v = T(0);
            const int r1 = TT.fq_boff[bk + 1];
            int r = TT.fq_boff[bk] + lane;
            int4 kn = r < r1 ? k4[r] : int4{-1, -1, -1, -1};
            T cn = r < r1 ? c[r] : T(0);
            for (; r < r1; r += 64) {
                const int4 kk = kn;
                const T c0 = cn;
                if (r + 64 < r1) kn = k4[r + 64], cn = c[r + 64];
                const T pr = prod_keys<T, NV>(a, kk, one).re;
                if (TT.fv_skip && fabs(pr) <= T(1e-20)) continue;
                v += c0 * pr;
            }
            if (f >= 0) xo[f] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// LDS carve-up of xrows_kernel (identical on host and device)
// ---------------------------------------------------------------------------
struct XRowsLds {
    size_t tile, wv, xneed, nk, zr, waves, wave_stride, qoff, total;
    __host__ __device__ XRowsLds(const XRowsParams& P, int elem) {
        auto al = [](size_t x) { return (x + 15) & ~size_t(15); };
        size_t o = 0;
        tile = o;
        o = al(o + (P.gtile ? 0 : (size_t)P.rows_max * P.CB * elem));
        wv = o;
        o = al(o + (size_t)P.wz_max * 2 * elem);
        xneed = o;
        o = al(o + (size_t)P.need_max * elem);
        nk = o;
        o = al(o + (size_t)P.wz_max * 4);
        zr = o;
        o = al(o + (size_t)P.need_max * (P.ntp + 1) * 4);
        waves = o;
        qoff = al((size_t)P.slot * elem);
        wave_stride = al(qoff + (size_t)(64 / P.CB) * P.need_max * 3 * elem);
        total = o + (size_t)P.NW * wave_stride;
    }
};

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// one nonzero of W': key (packed), weight W. The four increments are formed
// first and added once, so the accumulators stay in registers.
template <class T>
__device__ __forceinline__ void q_term(const T* rec, const PairRec& R, int pk, Cpx<T> W, T& ry,
                                       T& a0, T& a1, T& a2) {
    const int key = (pk >> 8) - 1;
    T d0 = W.re, d1 = T(0), d2 = T(0), d3 = T(0);
    if (key >= 0) {
        const T* y = rec + R.ylm(key);
        d0 = W.re * y[0] - W.im * y[1];
        d1 = W.re * y[2] - W.im * y[3];
        d2 = W.re * y[4] - W.im * y[5];
        d3 = W.re * y[6] - W.im * y[7];
    }
    ry += d0;
    a0 += d1;
    a1 += d2;
    a2 += d3;
}

// d x_f / d d for one pair record: W' nonzeros z in [za, zb) of feature f
// restricted to the pair's type pair. Every nonzero of one linear feature has
// the feature's radial index, so f_n and f_n' are read once; f_n < 1e-20
// drops the pair (local.cpp:164).
template <class T>
__device__ __forceinline__ void q_feature(const T* rec, const PairRec& R,
                                          const Cpx<T>* __restrict__ Wv,
                                          const int* __restrict__ nk, int za, int zb, T& q0,
                                          T& q1, T& q2) {
    q0 = q1 = q2 = T(0);
    if (za >= zb) return;
    const int nid = nk[za] & 255;
    const T fv = rec[R.fn() + nid];
    if (fv < T(1e-20)) return;
    const T fdv = rec[R.fd() + nid];
    const T ir = T(1) / rec[3];
    const T ux = rec[0] * ir, uy = rec[1] * ir, uz = rec[2] * ir;
    T a0 = T(0), a1 = T(0), a2 = T(0), ry = T(0);
    // four nonzeros per step: their keys and weights are loaded together
    int z = za;
    for (; z + 4 <= zb; z += 4) {
        const int k0 = nk[z], k1 = nk[z + 1], k2 = nk[z + 2], k3 = nk[z + 3];
        const Cpx<T> w0 = Wv[z], w1 = Wv[z + 1], w2 = Wv[z + 2], w3 = Wv[z + 3];
        q_term<T>(rec, R, k0, w0, ry, a0, a1, a2);
        q_term<T>(rec, R, k1, w1, ry, a0, a1, a2);
        q_term<T>(rec, R, k2, w2, ry, a0, a1, a2);
        q_term<T>(rec, R, k3, w3, ry, a0, a1, a2);
    }
    for (; z < zb; ++z) q_term<T>(rec, R, nk[z], Wv[z], ry, a0, a1, a2);
    ry *= fdv;
    q0 = ry * ux + fv * a0;
    q1 = ry * uy + fv * a1;
    q2 = ry * uz + fv * a2;
}

// G: the row tile is the structure's own block of X in global memory
// (structures whose 7 + 3N rows do not fit LDS); its adds are device atomics.
template <class T, bool G>
__global__ void __launch_bounds__(sizeof(T) == 8 ? 512 : 1024) xrows_kernel(ModelDev md, ColumnsDev cd, BatchDev bd,
                                                    BlockWz bw, XRowsParams P,
                                                    const Cpx<T>* __restrict__ wall, int w_stride,
                                                    const T* __restrict__ xall, int x_stride,
                                                    int atom_base,
                                                    const T* __restrict__ pd, PairRec R,
                                                    T* __restrict__ X,
                                                    const double* __restrict__ roww,
                                                    double* __restrict__ xe_sum,
                                                    double* __restrict__ xe_sq) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    int blk, s;
    if (P.xcd_map) {
        // the blocks of one structure run on one XCD (workgroups are dealt to
        // the 8 XCDs round robin), so its pair records are shared in one L2
        const int xcd = blockIdx.x & 7, k = blockIdx.x >> 3;
        s = P.st0 + (k / cd.n_blocks) * 8 + xcd;
        blk = k % cd.n_blocks;
        if (s >= P.st_end) return;
    } else {
        blk = blockIdx.x % cd.n_blocks;
        s = P.st0 + blockIdx.x / cd.n_blocks;
    }
    const XRowsLds Lo(P, sizeof(T));
    Cpx<T>* Wv = reinterpret_cast<Cpx<T>*>(smem_raw + Lo.wv);
    T* xneed = reinterpret_cast<T*>(smem_raw + Lo.xneed);
    int* nk = reinterpret_cast<int*>(smem_raw + Lo.nk);
    int* zr = reinterpret_cast<int*>(smem_raw + Lo.zr);

    const int c0 = cd.blk_c0[blk];
    const int ncol = cd.blk_c0[blk + 1] - c0;
    const int rbase = bd.st_row_off[s] - P.row_base;
    T* tile = G ? X + (size_t)rbase * P.ld + c0 : reinterpret_cast<T*>(smem_raw + Lo.tile);
    const int ts = G ? P.ld : P.CB;  // row stride of the tile
    const int a0 = bd.st_atom_off[s], nat = bd.st_atom_off[s + 1] - a0;
    const int rows = 7 + 3 * nat;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int NW = blockDim.x >> 6;
    const int ntp = md.n_tp;
    const int CB = P.CB, PPW = 64 / CB;
    const int col = lane % CB, sub = lane / CB;
    const bool linear = cd.blk_kind[blk] == 1;
    const int pre = R.prefix(cd.blk_nkey[blk]);
    T* slot = reinterpret_cast<T*>(smem_raw + Lo.waves + (size_t)wave * Lo.wave_stride);
    T* qb = reinterpret_cast<T*>(smem_raw + Lo.waves + (size_t)wave * Lo.wave_stride + Lo.qoff);
    const int nb1 = cd.n_blocks + 1;
    const int c = c0 + col;
    const bool col_ok = col < ncol;
    const int nfac = col_ok ? cd.nfac[c] : 0;

    for (int e = tid; e < rows * CB; e += blockDim.x)
        if (!G || e % CB < ncol) tile[(size_t)(e / CB) * ts + e % CB] = T(0);
    if (G) __threadfence();

    // this lane's first element of a staged group of records: record p0s, offset o0s
    const int p0s = lane / pre, o0s = lane % pre;

    T sacc[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
    T eacc = T(0);

    for (int ia = 0; ia < nat; ++ia) {
        const int i = a0 + ia;
        const int t1 = bd.types[i];
        const Cpx<T>* wi = wall + (size_t)(i - atom_base) * w_stride;
        const T* xi = xall + (size_t)(i - atom_base) * x_stride;
        const int nf0 = cd.need_off[t1 * nb1 + blk];
        const int nneed = cd.need_off[t1 * nb1 + blk + 1] - nf0;
        const int* need = cd.need_feat + nf0;
        const int z0 = bw.off[t1 * nb1 + blk];
        const int nz = bw.off[t1 * nb1 + blk + 1] - z0;
        __syncthreads();
        for (int k = tid; k < nneed; k += blockDim.x) xneed[k] = xi[need[k]];
        for (int z = tid; z < nz; z += blockDim.x) {
            Wv[z] = wi[bw.nz[z0 + z]];
            nk[z] = bw.nk[z0 + z];
        }
        const int* zsrc = bw.zr + (size_t)nf0 * (ntp + 1);
        for (int e = tid; e < nneed * (ntp + 1); e += blockDim.x) zr[e] = zsrc[e];
        __syncthreads();

        // this lane's column: need indices of its factors and coefficients
        int cn0 = -1, cn1 = -1, cn2 = -1;
        T k0 = T(0), k1 = T(0), k2 = T(0);
        if (col_ok) {
            const int* cn = cd.col_need + ((size_t)t1 * cd.n_cols + c) * 3;
            cn0 = cn[0];
            if (nfac > 1) cn1 = cn[1];
            if (nfac > 2) cn2 = cn[2];
        }
        const bool active = cn0 >= 0;
        if (active) {
            T e = T(0);
            if (nfac == 1) {
                e = xneed[cn0], k0 = T(1);
            } else if (nfac == 2) {
                const T x0 = xneed[cn0], x1 = xneed[cn1];
                e = x0 * x1, k0 = x1, k1 = x0;
            } else {
                const T x0 = xneed[cn0], x1 = xneed[cn1], x2 = xneed[cn2];
                e = x0 * x1 * x2, k0 = x1 * x2, k1 = x0 * x2, k2 = x0 * x1;
            }
            if (wave == 0 && sub == 0) eacc += e;
        }
        T self0 = T(0), self1 = T(0), self2 = T(0);
        const int p0 = bd.nb_off[i], p1 = bd.nb_off[i + 1];

        // records of a group of PPW consecutive pairs are staged through
        // registers: the next group is loaded while the current one is used
        constexpr int kStageRegs = StageRegs<T>::n;
        T stage[kStageRegs];
        auto load_group = [&](int pc) {
            const int nel = min(PPW, p1 - pc) * pre;
            const T* src = pd + (size_t)(pc - P.pair_base) * R.stride;
            int p = p0s, o = o0s;
#pragma unroll
            for (int r = 0; r < kStageRegs; ++r) {
                if (lane + 64 * r < nel) stage[r] = src[(size_t)p * R.stride + o];
                o += 64;
                while (o >= pre) o -= pre, ++p;
            }
        };
        int pc = p0 + wave * PPW;
        if (pc < p1) load_group(pc);
        for (; pc < p1; pc += NW * PPW) {
            const int np = min(PPW, p1 - pc);
#pragma unroll
            for (int r = 0; r < kStageRegs; ++r)
                if (lane + 64 * r < np * pre) slot[lane + 64 * r] = stage[r];
            const int nxt = pc + NW * PPW;
            if (nxt < p1) load_group(nxt);
            wave_sync();
            const T* rec = slot + (size_t)sub * pre;
            const bool pv = sub < np && rec[4] != T(0);
            if (linear) {
                if (pv && active) {
                    const int tp = int(rec[5]);
                    T v0, v1, v2;
                    q_feature<T>(rec, R, Wv, nk, zr[cn0 * (ntp + 1) + tp],
                                 zr[cn0 * (ntp + 1) + tp + 1], v0, v1, v2);
                    self0 += v0, self1 += v1, self2 += v2;
                    T* rj = tile + (size_t)(7 + 3 * int(rec[6])) * ts + col;
                    atomicAdd(rj, -v0);
                    atomicAdd(rj + ts, -v1);
                    atomicAdd(rj + 2 * ts, -v2);
                    const T dx = rec[0], dy = rec[1], dz = rec[2];
                    sacc[0] -= v0 * dx, sacc[1] -= v1 * dy, sacc[2] -= v2 * dz;
                    sacc[3] -= v0 * dy, sacc[4] -= v1 * dz, sacc[5] -= v2 * dx;
                }
            } else {
                if (pv) {
                    const int tp = int(rec[5]);
                    for (int k = col; k < nneed; k += CB) {
                        T q0, q1, q2;
                        q_feature<T>(rec, R, Wv, nk, zr[k * (ntp + 1) + tp],
                                     zr[k * (ntp + 1) + tp + 1], q0, q1, q2);
                        T* qq = qb + ((size_t)sub * P.need_max + k) * 3;
                        qq[0] = q0, qq[1] = q1, qq[2] = q2;
                    }
                }
                wave_sync();
                if (pv && active) {
                    const T* q = qb + (size_t)sub * P.need_max * 3;
                    T v0 = k0 * q[3 * cn0], v1 = k0 * q[3 * cn0 + 1], v2 = k0 * q[3 * cn0 + 2];
                    if (nfac > 1) {
                        v0 += k1 * q[3 * cn1], v1 += k1 * q[3 * cn1 + 1], v2 += k1 * q[3 * cn1 + 2];
                    }
                    if (nfac > 2) {
                        v0 += k2 * q[3 * cn2], v1 += k2 * q[3 * cn2 + 1], v2 += k2 * q[3 * cn2 + 2];
                    }
                    self0 += v0, self1 += v1, self2 += v2;
                    T* rj = tile + (size_t)(7 + 3 * int(rec[6])) * ts + col;
                    atomicAdd(rj, -v0);
                    atomicAdd(rj + ts, -v1);
                    atomicAdd(rj + 2 * ts, -v2);
                    const T dx = rec[0], dy = rec[1], dz = rec[2];
                    sacc[0] -= v0 * dx, sacc[1] -= v1 * dy, sacc[2] -= v2 * dz;
                    sacc[3] -= v0 * dy, sacc[4] -= v1 * dz, sacc[5] -= v2 * dx;
                }
            }
            wave_sync();
        }
        if (active) {
            T* ri = tile + (size_t)(7 + 3 * ia) * ts + col;
            atomicAdd(ri, self0);
            atomicAdd(ri + ts, self1);
            atomicAdd(ri + 2 * ts, self2);
        }
    }
    if (col_ok) {
        for (int k = 0; k < 6; ++k) atomicAdd(tile + (size_t)(1 + k) * ts + col, sacc[k]);
        if (wave == 0 && sub == 0) atomicAdd(tile + col, eacc);
    }
    if (G) __threadfence();
    __syncthreads();
    // write out (weights applied per row); the raw energy row feeds the scale sums
    for (int e = tid; e < rows * CB; e += blockDim.x) {
        const int r = e / CB, cc = e % CB;
        if (cc >= ncol) continue;
        const T v = G ? __hip_atomic_load(tile + (size_t)r * ts + cc, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT)
                      : tile[e];
        const double w = roww ? roww[rbase + r] : 1.0;
        X[(size_t)(rbase + r) * P.ld + c0 + cc] = T(w * double(v));
        if (r == 0 && xe_sum) {
            atomicAdd(xe_sum + c0 + cc, double(v));
            atomicAdd(xe_sq + c0 + cc, double(v) * double(v));
        }
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
size_t pairdata_smem(const ModelDev& md, int elem) {
    const PairLayout<double> L(md.n_fn, md.n_lmh);
    return (size_t)kPdPairs * L.stride * elem;
}

void launch_pairdata(hipStream_t st, bool fp64, const ModelDev& md, const BatchDev& bd, int atom0,
                     int n_atoms, long pair_base, void* pd, const PairRec& R, int lmax, int* tpo) {
    if (tpo && md.n_tp > kMaxPairTypes) throw std::invalid_argument("too many type pairs");
    const size_t sm = pairdata_smem(md, fp64 ? 8 : 4);
    if (fp64) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&pairdata_kernel<double>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
        pairdata_kernel<double><<<n_atoms, 64, sm, st>>>(md, bd, atom0, n_atoms, pair_base,
                                                         reinterpret_cast<double*>(pd), R, lmax, tpo);
    } else {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&pairdata_kernel<float>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
        pairdata_kernel<float><<<n_atoms, 64, sm, st>>>(md, bd, atom0, n_atoms, pair_base,
                                                        reinterpret_cast<float*>(pd), R, lmax, tpo);
    }
}

size_t xrows_smem(const XRowsParams& P, int elem) { return XRowsLds(P, elem).total; }

void launch_anlm_rec(hipStream_t st, bool fp64, const ModelDev& md, const BatchDev& bd, int atom0,
                     int n_atoms, long pair_base, const void* pd, const PairRec& R,
                     const int* tpo, void* a_full, int a_stride, int max_noconj) {
    const int elem = fp64 ? 8 : 4;
    const size_t sm = (size_t)kArPairs * (md.n_fn + 2 * md.n_lmh) * elem + (size_t)max_noconj * 2 * elem;
    if (fp64) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&anlm_rec_kernel<double>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
        anlm_rec_kernel<double><<<n_atoms, 256, sm, st>>>(
            md, bd, atom0, n_atoms, pair_base, reinterpret_cast<const double*>(pd), R, tpo,
            reinterpret_cast<Cpx<double>*>(a_full), a_stride);
    } else {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&anlm_rec_kernel<float>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
        anlm_rec_kernel<float><<<n_atoms, 256, sm, st>>>(
            md, bd, atom0, n_atoms, pair_base, reinterpret_cast<const float*>(pd), R, tpo,
            reinterpret_cast<Cpx<float>*>(a_full), a_stride);
    }
}

int xrows_stage_limit(bool fp64) {
    return 64 * (fp64 ? StageRegs<double>::n : StageRegs<float>::n);
}

int xrows_max_waves(bool fp64) { return fp64 ? 8 : 16; }


template <class T, int NF, int NV>
static void launch_wprime_t(hipStream_t st, const ModelDev& md, const BatchDev& bd,
                            const void* a_full, int a_stride, int atom0, int n_atoms, void* w,
                            int w_stride, void* x, int x_stride, int max_nlm) {
    const size_t sm = (size_t)(max_nlm + 1) * 2 * sizeof(T);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&wprime_kernel<T, NF, NV>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    wprime_kernel<T, NF, NV><<<n_atoms, 256, sm, st>>>(
        md, bd, reinterpret_cast<const Cpx<T>*>(a_full), a_stride, atom0, n_atoms,
        reinterpret_cast<Cpx<T>*>(w), w_stride, reinterpret_cast<T*>(x), x_stride);
}

void launch_wprime(hipStream_t st, bool fp64, const ModelDev& md, const BatchDev& bd,
                   const void* a_full, int a_stride, int atom0, int n_atoms, void* w, int w_stride,
                   void* x, int x_stride, int max_nlm, int w_maxf, int v_maxf) {
    if (w_maxf > 3 || v_maxf > 4) throw std::invalid_argument("invariant order above 4");
    // instantiations: W' terms with <= 2 or 3 factors, value terms with <= 3 or 4
    const int nf = w_maxf <= 2 ? 2 : 3, nv = v_maxf <= 3 ? 3 : 4;
#define PLM_WP(T, NF, NV)                                                                       \
    launch_wprime_t<T, NF, NV>(st, md, bd, a_full, a_stride, atom0, n_atoms, w, w_stride, x, \
                               x_stride, max_nlm)
    if (fp64) {
        if (nf == 2 && nv == 3) PLM_WP(double, 2, 3);
        else if (nf == 2) PLM_WP(double, 2, 4);
        else if (nv == 3) PLM_WP(double, 3, 3);
        else PLM_WP(double, 3, 4);
    } else {
        if (nf == 2 && nv == 3) PLM_WP(float, 2, 3);
        else if (nf == 2) PLM_WP(float, 2, 4);
        else if (nv == 3) PLM_WP(float, 3, 3);
        else PLM_WP(float, 3, 4);
    }
#undef PLM_WP
}

template <class T>
static void launch_xrows_t(hipStream_t st, const ModelDev& md, const ColumnsDev& cd,
                           const BatchDev& bd, const BlockWz& bw, const XRowsParams& P,
                           const void* w, int w_stride, const void* x, int x_stride,
                           int atom_base, const void* pd, const PairRec& R, void* X,
                           const double* roww, double* xe_sum, double* xe_sq) {
    const size_t sm = xrows_smem(P, sizeof(T));
    const auto kern = P.gtile ? &xrows_kernel<T, true> : &xrows_kernel<T, false>;
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    const int n_st = P.st_end - P.st0;
    const long grid = P.xcd_map ? (long)((n_st + 7) / 8) * 8 * cd.n_blocks : (long)n_st * cd.n_blocks;
    kern<<<grid, 64 * P.NW, sm, st>>>(
        md, cd, bd, bw, P, reinterpret_cast<const Cpx<T>*>(w), w_stride,
        reinterpret_cast<const T*>(x), x_stride, atom_base, reinterpret_cast<const T*>(pd), R,
        reinterpret_cast<T*>(X), roww, xe_sum, xe_sq);
}

void launch_xrows(hipStream_t st, bool fp64, const ModelDev& md, const ColumnsDev& cd,
                  const BatchDev& bd, const BlockWz& bw, const XRowsParams& P, const void* w,
                  int w_stride, const void* x, int x_stride, int atom_base, const void* pd,
                  const PairRec& R, void* X, const double* roww, double* xe_sum, double* xe_sq) {
    if (fp64)
        launch_xrows_t<double>(st, md, cd, bd, bw, P, w, w_stride, x, x_stride, atom_base, pd, R,
                               X, roww, xe_sum, xe_sq);
    else
        launch_xrows_t<float>(st, md, cd, bd, bw, P, w, w_stride, x, x_stride, atom_base, pd, R,
                              X, roww, xe_sum, xe_sq);
}

}  // namespace plm
