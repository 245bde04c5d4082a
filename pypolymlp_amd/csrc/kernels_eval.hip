// HIP kernels of PolymlpEval (energy, forces, stress) for gfx950.
//
// Reference loop nests:
//   PolymlpEval::compute_sum_of_prod_anlmtp   src/cxx/src/compute/polymlp_eval.cpp:337-404
//   Potential::compute_sum_of_prod_anlmtp     src/cxx/src/polymlp/polymlp_potential.cpp:235-321
//   PolymlpEval::eval_gtinv (half-pair loop)  polymlp_eval.cpp:182-291
//   PolymlpEval::eval_pair / compute_sum_of_prod_antp  polymlp_eval.cpp:46-180,
//                                             polymlp_potential.cpp:203-233
//   collect_properties                        polymlp_eval.cpp:293-335
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>

#include "device_math.hpp"
#include "device_tables.hpp"
#include "eval_tables.hpp"
#include "kernels.hpp"

namespace plm {

template <class T>
__device__ __forceinline__ Cpx<T> prod_keys3(const Cpx<T>* a, const int* keys) {
    Cpx<T> v = {T(1), T(0)};
    for (int k = 0; k < 3; ++k) {
        const int id = keys[k];
        if (id < 0) break;
        v = cmul(v, a[id]);
    }
    return v;
}

// ---------------------------------------------------------------------------
// Per atom: features, products of features, and
//   prod_sum_{e,f}[h] = sum_terms coeff_{e,f} * prod_features * prod_anlm
// One 64-lane workgroup per atom; the atom's order parameters and features
// staged in LDS (the term products gather from them at random keys).
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(64) prodsum_kernel(ModelDev md, EvalDev ed, BatchDev bd,
                                                     const Cpx<T>* __restrict__ a_full,
                                                     int a_stride, Cpx<T>* __restrict__ ps_e,
                                                     Cpx<T>* __restrict__ ps_f, int ps_stride,
                                                     int pair_model) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    // the atom's order parameters (a_stride complex values), then its features
    Cpx<T>* a = reinterpret_cast<Cpx<T>*>(smem_raw);
    T* xs = reinterpret_cast<T*>(a + a_stride);
    const int i = blockIdx.x;
    if (i >= bd.n_atoms_total) return;
    const int t1 = bd.types[i];
    const TypeDev& TT = md.t[t1];
    const EvalTypeDev& E = ed.t[t1];
    const Cpx<T>* ag = a_full + (size_t)i * a_stride;
    for (int k = threadIdx.x; k < TT.n_nlm; k += blockDim.x) a[k] = ag[k];
    __syncthreads();
    for (int f = threadIdx.x; f < TT.n_lin; f += blockDim.x) {
        T v = T(0);
        for (int e = TT.fv_off[f]; e < TT.fv_off[f + 1]; ++e) {
            Cpx<T> pr = {T(1), T(0)};
            const int* keys = TT.fv_keys + 4 * (size_t)e;
            for (int k = 0; k < 4 && keys[k] >= 0; ++k) pr = cmul(pr, a[keys[k]]);
            if (TT.fv_skip && fabs(pr.re) <= T(1e-20)) continue;
            v += T(TT.fv_coeff[e]) * pr.re;
        }
        xs[f] = v;
    }
    __syncthreads();
    Cpx<T>* oe = ps_e + (size_t)i * ps_stride;
    Cpx<T>* of = ps_f + (size_t)i * ps_stride;
    for (int h = threadIdx.x; h < E.n_heads; h += blockDim.x) {
        T er = T(0), ei = T(0), fr = T(0), fi = T(0);
        for (int e = E.pot_off[h]; e < E.pot_off[h + 1]; ++e) {
            const int* pk = E.pot_pf + 2 * (size_t)e;
            T fval = T(1);
            if (pk[0] >= 0) fval *= xs[pk[0]];
            if (pk[1] >= 0) fval *= xs[pk[1]];
            const Cpx<T> pr = prod_keys3<T>(a, E.pot_pd + 3 * (size_t)e);
            const T se = fval * T(E.pot_ce[e]), sf = fval * T(E.pot_cf[e]);
            er += pr.re * se, ei += pr.im * se;
            fr += pr.re * sf, fi += pr.im * sf;
        }
        if (pair_model) er *= T(0.5), ei *= T(0.5), fr *= T(0.5), fi *= T(0.5);
        oe[h] = {er, ei};
        of[h] = {fr, fi};
    }
}

// ---------------------------------------------------------------------------
// Half-pair loop: one lane per pair, radial + Y_lm (+ gradients) in registers
// bounded by LB (max l) and NB (radial functions).
// ---------------------------------------------------------------------------
template <class T, int LB, int NB>
__global__ void __launch_bounds__(64) pair_kernel(ModelDev md, EvalDev ed, BatchDev bd,
                                                  const int* __restrict__ atom_st,
                                                  const int* __restrict__ hb_off,
                                                  const int* __restrict__ hb_j,
                                                  const double* __restrict__ hb_d,
                                                  const Cpx<T>* __restrict__ ps_e,
                                                  const Cpx<T>* __restrict__ ps_f, int ps_stride,
                                                  double* __restrict__ energy,
                                                  double* __restrict__ forces,
                                                  double* __restrict__ stress, int pair_model) {
    constexpr int NLMH = (LB + 1) * (LB + 2) / 2;
    const int i = blockIdx.x;
    if (i >= bd.n_atoms_total) return;
    const int t1 = bd.types[i];
    const EvalTypeDev& E = ed.t[t1];
    const int p0 = hb_off[i], p1 = hb_off[i + 1];
    double e_acc = 0.0, s_acc[6] = {0, 0, 0, 0, 0, 0};
    double fi_acc[3] = {0, 0, 0};
    const Cpx<T>* Pi_e = ps_e + (size_t)i * ps_stride;
    const Cpx<T>* Pi_f = ps_f + (size_t)i * ps_stride;
    for (int k = p0 + threadIdx.x; k < p1; k += blockDim.x) {
        const int j = hb_j[k];
        // diff_ij: r_i - r_j - T
        const double dx = -hb_d[3 * (size_t)k], dy = -hb_d[3 * (size_t)k + 1],
                     dz = -hb_d[3 * (size_t)k + 2];
        const double dis = sqrt(dx * dx + dy * dy + dz * dz);
        if (dis >= md.cutoff) continue;
        const int tp = md.tp_of[t1 * md.n_type + bd.types[j]];
        T fn[NB], fd[NB];
        radial<T>(T(dis), T(md.cutoff), md.rad + (size_t)tp * md.n_fn * 2, md.rad_count[tp], fn, fd);
        const Cpx<T>* Pj_e = ps_e + (size_t)j * ps_stride;
        const Cpx<T>* Pj_f = ps_f + (size_t)j * ps_stride;
        double eij = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
        const int h0 = E.hd_off[tp], h1 = E.hd_off[tp + 1];
        if (pair_model) {
            T ef = T(0), ff = T(0);
            for (int h = h0; h < h1; ++h) {
                const int nid = E.hd_nid[h];
                const int hi = E.hd_i[h], hj = E.hd_j[h];
                ef += fn[nid] * (Pi_e[hi].re + Pj_e[hj].re);
                ff += fd[nid] * (Pi_f[hi].re + Pj_f[hj].re);
            }
            eij = double(ef);
            const double fij = -double(ff) / dis;
            fx = fij * dx, fy = fij * dy, fz = fij * dz;
        } else {
            Cpx<T> Y[NLMH], dY[3 * NLMH];
            T pbuf[NLMH], qbuf[NLMH];
            ylm<T>(md.max_l, T(dx), T(dy), T(dz), T(dis), Y, dY, pbuf, qbuf);
            const T ir = T(1) / T(dis);
            T ea = T(0), fxa = T(0), fya = T(0), fza = T(0);
            for (int h = h0; h < h1; ++h) {
                const int nid = E.hd_nid[h];
                const T fv = fn[nid];
                if (fv < T(1e-20)) continue;
                const int key = E.hd_ylm[h];
                const int hi = E.hd_i[h], hj = E.hd_j[h];
                const T sj = T(E.hd_signj[h]);
                const T w = E.hd_m0[h] ? T(0.5) : T(1);
                const Cpx<T> se = {(Pi_e[hi].re + Pj_e[hj].re * sj) * w,
                                   (Pi_e[hi].im + Pj_e[hj].im * sj) * w};
                const Cpx<T> sf = {(Pi_f[hi].re + Pj_f[hj].re * sj) * w,
                                   (Pi_f[hi].im + Pj_f[hj].im * sj) * w};
                const Cpx<T> y = Y[key];
                ea += fv * (y.re * se.re - y.im * se.im);
                const T d1r = fd[nid] * y.re * ir, d1i = fd[nid] * y.im * ir;
                // val_a = -(d1 * d_a + fn * dY_a)
                const Cpx<T> vx = {-(d1r * T(dx) + fv * dY[3 * key].re), -(d1i * T(dx) + fv * dY[3 * key].im)};
                const Cpx<T> vy = {-(d1r * T(dy) + fv * dY[3 * key + 1].re), -(d1i * T(dy) + fv * dY[3 * key + 1].im)};
                const Cpx<T> vz = {-(d1r * T(dz) + fv * dY[3 * key + 2].re), -(d1i * T(dz) + fv * dY[3 * key + 2].im)};
                fxa += re_mul(vx, sf);
                fya += re_mul(vy, sf);
                fza += re_mul(vz, sf);
            }
            eij = double(ea), fx = double(fxa), fy = double(fya), fz = double(fza);
        }
        e_acc += eij;
        fi_acc[0] += fx, fi_acc[1] += fy, fi_acc[2] += fz;
        atomicAdd(forces + 3 * (size_t)j, -fx);
        atomicAdd(forces + 3 * (size_t)j + 1, -fy);
        atomicAdd(forces + 3 * (size_t)j + 2, -fz);
        s_acc[0] += dx * fx, s_acc[1] += dy * fy, s_acc[2] += dz * fz;
        s_acc[3] += dx * fy, s_acc[4] += dy * fz, s_acc[5] += dz * fx;
    }
    // wave reduction (64 lanes)
    for (int off = 32; off > 0; off >>= 1) {
        e_acc += __shfl_down(e_acc, off);
        for (int a = 0; a < 3; ++a) fi_acc[a] += __shfl_down(fi_acc[a], off);
        for (int a = 0; a < 6; ++a) s_acc[a] += __shfl_down(s_acc[a], off);
    }
    if (threadIdx.x == 0) {
        const int s = atom_st[i];
        atomicAdd(energy + s, e_acc);
        for (int a = 0; a < 3; ++a) atomicAdd(forces + 3 * (size_t)i + a, fi_acc[a]);
        for (int a = 0; a < 6; ++a) atomicAdd(stress + 6 * (size_t)s + a, s_acc[a]);
    }
}

template <class T>
void launch_prodsum_t(hipStream_t st, const ModelDev& md, const EvalDev& ed, const BatchDev& bd,
                      const void* a_full, int a_stride, void* ps_e, void* ps_f, int ps_stride,
                      int max_lin, int pair_model) {
    const size_t sm = (size_t)a_stride * 2 * sizeof(T) + (size_t)std::max(1, max_lin) * sizeof(T);
    if (sm > 160 * 1024) throw std::invalid_argument("per-atom evaluation data exceed LDS");
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&prodsum_kernel<T>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    prodsum_kernel<T><<<bd.n_atoms_total, 64, sm, st>>>(
        md, ed, bd, reinterpret_cast<const Cpx<T>*>(a_full), a_stride,
        reinterpret_cast<Cpx<T>*>(ps_e), reinterpret_cast<Cpx<T>*>(ps_f), ps_stride, pair_model);
}

void launch_prodsum(hipStream_t st, bool fp64, const ModelDev& md, const EvalDev& ed,
                    const BatchDev& bd, const void* a_full, int a_stride, void* ps_e, void* ps_f,
                    int ps_stride, int max_lin, int pair_model) {
    if (fp64)
        launch_prodsum_t<double>(st, md, ed, bd, a_full, a_stride, ps_e, ps_f, ps_stride, max_lin,
                                 pair_model);
    else
        launch_prodsum_t<float>(st, md, ed, bd, a_full, a_stride, ps_e, ps_f, ps_stride, max_lin,
                                pair_model);
}

template <class T, int LB, int NB>
void launch_pair_lb(hipStream_t st, const ModelDev& md, const EvalDev& ed, const BatchDev& bd,
                    const int* atom_st, const int* hb_off, const int* hb_j, const double* hb_d,
                    const void* ps_e, const void* ps_f, int ps_stride, double* energy,
                    double* forces, double* stress, int pair_model) {
    pair_kernel<T, LB, NB><<<bd.n_atoms_total, 64, 0, st>>>(
        md, ed, bd, atom_st, hb_off, hb_j, hb_d, reinterpret_cast<const Cpx<T>*>(ps_e),
        reinterpret_cast<const Cpx<T>*>(ps_f), ps_stride, energy, forces, stress, pair_model);
}

template <class T>
void launch_pair_t(hipStream_t st, const ModelDev& md, const EvalDev& ed, const BatchDev& bd,
                   const int* atom_st, const int* hb_off, const int* hb_j, const double* hb_d,
                   const void* ps_e, const void* ps_f, int ps_stride, double* energy,
                   double* forces, double* stress, int pair_model) {
    const int L = md.max_l;
    if (md.n_fn > 32) throw std::invalid_argument("evaluation supports at most 32 radial functions");
    if (L <= 4)
        launch_pair_lb<T, 4, 32>(st, md, ed, bd, atom_st, hb_off, hb_j, hb_d, ps_e, ps_f, ps_stride,
                                 energy, forces, stress, pair_model);
    else if (L <= 8)
        launch_pair_lb<T, 8, 32>(st, md, ed, bd, atom_st, hb_off, hb_j, hb_d, ps_e, ps_f, ps_stride,
                                 energy, forces, stress, pair_model);
    else if (L <= 12)
        launch_pair_lb<T, 12, 32>(st, md, ed, bd, atom_st, hb_off, hb_j, hb_d, ps_e, ps_f, ps_stride,
                                  energy, forces, stress, pair_model);
    else
        throw std::invalid_argument("evaluation supports max_l <= 12");
}

void launch_pair(hipStream_t st, bool fp64, const ModelDev& md, const EvalDev& ed,
                 const BatchDev& bd, const int* atom_st, const int* hb_off, const int* hb_j,
                 const double* hb_d, const void* ps_e, const void* ps_f, int ps_stride,
                 double* energy, double* forces, double* stress, int pair_model) {
    if (fp64)
        launch_pair_t<double>(st, md, ed, bd, atom_st, hb_off, hb_j, hb_d, ps_e, ps_f, ps_stride,
                              energy, forces, stress, pair_model);
    else
        launch_pair_t<float>(st, md, ed, bd, atom_st, hb_off, hb_j, hb_d, ps_e, ps_f, ps_stride,
                             energy, forces, stress, pair_model);
}

__global__ void __launch_bounds__(256) fill_zero_kernel(double* __restrict__ x, size_t n) {
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (size_t)gridDim.x * blockDim.x)
        x[k] = 0.0;
}

void launch_fill_zero(hipStream_t st, double* x, size_t n) {
    if (n == 0) return;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 4096);
    fill_zero_kernel<<<(unsigned)blocks, 256, 0, st>>>(x, n);
}

}  // namespace plm
