// HIP kernels of the X path for gfx950: brute-force neighbour lists and the
// order parameters a_{nlm,tp}(i).
//
// Reference loop nests these kernels replace:
//   NeighborFull          src/cxx/src/compute/neighbor_full.cpp:10-76
//   NeighborHalf          src/cxx/src/compute/neighbor_half.cpp:10-100
//   Local::compute_anlmtp_d (values, conjugate fill)  local.cpp:116-217
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <stdexcept>

#include "device_math.hpp"
#include "device_tables.hpp"
#include "kernels.hpp"
#include "pair_data.hpp"

namespace plm {

void check_hip(hipError_t e, const char* what);

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
    const PairLayout<T> L(md.n_fn, md.n_lmh, false);  // values only
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

// Order parameters with register sums (evaluation, at most NT Q heads per
// type, NT threads per atom): thread t owns the entries t + NT q (q < Q) of the
// atom type's list
// grouped by type pair, holds their type pair, radial index, Y key and head
// in registers for the whole atom, and adds f_n Y_lm of every pair of its
// entries' type pair into a register sum; the sums go to their heads at the
// end. Pairs are taken in the same order as anlm_kernel, so the sums are the
// same; the entry tables are read once per atom instead of once per pair and
// the LDS read-modify-write per pair and entry is gone.
template <class T, int Q, int NT>
__global__ void __launch_bounds__(NT) anlm_reg_kernel(ModelDev md, BatchDev bd,
                                                      Cpx<T>* __restrict__ a_full, int a_stride,
                                                      int atom0, int n_atoms, int negate) {
    extern __shared__ unsigned char smem_raw[];
    const int ia = blockIdx.x;
    if (ia >= n_atoms) return;
    const int i = atom0 + ia;
    const int t1 = bd.types[i];
    const TypeDev& TT = md.t[t1];
    const PairLayout<T> L(md.n_fn, md.n_lmh, false);  // values only
    const int NPS = kAnlmPairs;
    T* pairs = reinterpret_cast<T*>(smem_raw);
    Cpx<T>* acc = reinterpret_cast<Cpx<T>*>(pairs + (size_t)NPS * L.stride);
    const int tid = threadIdx.x;
    const int n_nc = TT.nc_tp_off[md.n_tp];
    int etp[Q], enid[Q], ekey[Q], ehd[Q];
    Cpx<T> a[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int h = tid + NT * q;
        etp[q] = -1, enid[q] = 0, ekey[q] = 0, ehd[q] = 0;
        a[q] = {T(0), T(0)};
        if (h < n_nc) {
            int tp = 0;
            while (TT.nc_tp_off[tp + 1] <= h) ++tp;
            etp[q] = tp, enid[q] = TT.nc_tp_nid[h], ekey[q] = TT.nc_tp_ylm[h];
            ehd[q] = TT.nc_tp_head[h];
        }
    }
    for (int h = tid; h < TT.n_noconj; h += blockDim.x) acc[h] = {T(0), T(0)};
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
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                if (etp[q] != tp) continue;
                const T f = fn[enid[q]];
                if (f < T(1e-20)) continue;
                if (ekey[q] < 0) {
                    a[q].re += f;
                } else {
                    const Cpx<T> y = Y[ekey[q]];
                    a[q].re += f * y.re;
                    a[q].im += f * y.im;
                }
            }
        }
    }
    __syncthreads();  // the zeroed heads, before the entries' sums land on theirs
#pragma unroll
    for (int q = 0; q < Q; ++q)
        if (etp[q] >= 0) acc[ehd[q]] = a[q];
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
    const PairLayout<double> L(md.n_fn, md.n_lmh, false);
    return (size_t)nthreads * L.stride * elem + (size_t)max_noconj * 2 * elem;
}

template <class T, int Q, int NT>
void launch_anlm_reg(hipStream_t st, const ModelDev& md, const BatchDev& bd, void* a_full,
                     int a_stride, int atom0, int n_atoms, size_t sm, int negate) {
    check_hip(hipFuncSetAttribute(reinterpret_cast<const void*>(&anlm_reg_kernel<T, Q, NT>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm),
              "anlm shared memory");
    anlm_reg_kernel<T, Q, NT><<<n_atoms, NT, sm, st>>>(md, bd, reinterpret_cast<Cpx<T>*>(a_full),
                                                       a_stride, atom0, n_atoms, negate);
}

// register sums with 2-4 entries per thread (measured: 128 x 2 for 165 heads
// 0.96 ms per 32,000 atoms against 1.37 ms for 64 x 4 and 1.06 ms for 256 x 1;
// 512 x 3 for 1,350 heads 2.20 ms per 8,192 atoms against 2.73 ms for 256 x 8),
// up to 2,048 heads per type; LDS sums beyond
// PLM_ANLM_SHAPE=2x128|3x256|3x512|4x512|lds forces one form (an error where
// the type's heads do not fit it)
template <class T>
void launch_anlm_t(hipStream_t st, const ModelDev& md, const BatchDev& bd, void* a_full,
                   int a_stride, int atom0, int n_atoms, int max_noconj, int negate) {
    const int B = 64;
    const size_t sm = anlm_smem(md, sizeof(T), kAnlmPairs, max_noconj);
    if (const char* e = getenv("PLM_ANLM_SHAPE")) {
        auto forced = [&](int cap) {
            if (max_noconj > cap)
                throw std::invalid_argument("PLM_ANLM_SHAPE: the model's heads exceed this form");
        };
        if (!strcmp(e, "2x128")) {
            forced(2 * 128);
            launch_anlm_reg<T, 2, 128>(st, md, bd, a_full, a_stride, atom0, n_atoms, sm, negate);
            return;
        } else if (!strcmp(e, "3x256")) {
            forced(3 * 256);
            launch_anlm_reg<T, 3, 256>(st, md, bd, a_full, a_stride, atom0, n_atoms, sm, negate);
            return;
        } else if (!strcmp(e, "3x512")) {
            forced(3 * 512);
            launch_anlm_reg<T, 3, 512>(st, md, bd, a_full, a_stride, atom0, n_atoms, sm, negate);
            return;
        } else if (!strcmp(e, "4x512")) {
            forced(4 * 512);
            launch_anlm_reg<T, 4, 512>(st, md, bd, a_full, a_stride, atom0, n_atoms, sm, negate);
            return;
        } else if (strcmp(e, "lds") != 0) {
            throw std::invalid_argument("PLM_ANLM_SHAPE: 2x128, 3x256, 3x512, 4x512 or lds");
        }
        // lds: the LDS sums below
    } else if (max_noconj <= 2 * 128) {
        launch_anlm_reg<T, 2, 128>(st, md, bd, a_full, a_stride, atom0, n_atoms, sm, negate);
        return;
    } else if (max_noconj <= 3 * 256) {
        launch_anlm_reg<T, 3, 256>(st, md, bd, a_full, a_stride, atom0, n_atoms, sm, negate);
        return;
    } else if (max_noconj <= 3 * 512) {
        launch_anlm_reg<T, 3, 512>(st, md, bd, a_full, a_stride, atom0, n_atoms, sm, negate);
        return;
    } else if (max_noconj <= 4 * 512) {
        launch_anlm_reg<T, 4, 512>(st, md, bd, a_full, a_stride, atom0, n_atoms, sm, negate);
        return;
    }
    check_hip(hipFuncSetAttribute(reinterpret_cast<const void*>(&anlm_kernel<T>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm),
              "anlm shared memory");
    anlm_kernel<T><<<n_atoms, B, sm, st>>>(md, bd, reinterpret_cast<Cpx<T>*>(a_full), a_stride,
                                           atom0, n_atoms, negate);
}

void launch_anlm(hipStream_t st, bool fp64, const ModelDev& md, const BatchDev& bd, void* a_full,
                 int a_stride, int atom0, int n_atoms, int max_noconj, int negate) {
    if (fp64) launch_anlm_t<double>(st, md, bd, a_full, a_stride, atom0, n_atoms, max_noconj, negate);
    else launch_anlm_t<float>(st, md, bd, a_full, a_stride, atom0, n_atoms, max_noconj, negate);
}

}  // namespace plm
