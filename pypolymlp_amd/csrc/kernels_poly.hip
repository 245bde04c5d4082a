// Polynomial columns of X (products of 2 or 3 linear features, the
// "factors") as one GEMM per structure instead of per-pair column work.
//
// For a column c = prod_s x_{f_s} (Model::model_order2/3, model.cpp:207-267):
//   force/stress rows:  X[r, c] = sum_i sum_s D_{f_s}[r, i] * m_{c\s}(i)
//   energy row:         X[e, c] = sum_i prod_s x_{f_s}(i)
// where D_f[r, i] = d x_f(i) / d xi_r is the derivative of factor f at atom i
// (rows r: 6 stress, then 3N forces) and m_{c\s}(i) the product of the other
// factors at atom i. Each (slot s, column) pair is an entry e = (f, u) with
// f the slot's factor and u the monomial of the others; entries are sorted by
// slot and factor, and Y[r][e] = sum_i D[i][f][r] M[i][u] is one GEMM per
// (slot, factor) over that factor's entries (B: the gathered monomials).
// Every column is then the sum of its nfac entries of Y.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_math.hpp"
#include "device_tables.hpp"
#include "kernels.hpp"
#include "pair_data.hpp"

namespace plm {

// factor values and monomials per atom: XF[i][K], M[i][U]
template <class T>
__global__ void poly_factor_kernel(ModelDev md, PolyDev pd, BatchDev bd, int atom0, int n_atoms,
                                   const T* __restrict__ xall, int x_stride,
                                   T* __restrict__ XF, T* __restrict__ M) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    T* xf = reinterpret_cast<T*>(smem_raw);
    const int ia = blockIdx.x;
    if (ia >= n_atoms) return;
    const int t = bd.types[atom0 + ia];
    const T* xi = xall + (size_t)ia * x_stride;
    for (int k = threadIdx.x; k < pd.K; k += blockDim.x) {
        const int fl = pd.fac_local[t * pd.K + k];
        const T v = fl >= 0 ? xi[fl] : T(0);
        xf[k] = v;
        XF[(size_t)ia * pd.K + k] = v;
    }
    __syncthreads();
    for (int u = threadIdx.x; u < pd.U; u += blockDim.x) {
        const int a = pd.mono_a[u], b = pd.mono_b[u];
        M[(size_t)ia * pd.U + u] = b < 0 ? xf[a] : xf[a] * xf[b];
    }
}

// D[i][f][r] for the factors f in [k0, k0 + kc): one workgroup per atom, one
// thread per factor walking the atom's pairs (rows of neighbours j and of
// atom i and the 6 stress rows accumulate in LDS / registers without
// atomics), then a coalesced write of the atom's rows.
template <class T>
__global__ void __launch_bounds__(128) poly_deriv_kernel(ModelDev md, PolyDev pd, BatchDev bd,
                                                         int atom0, int n_atoms, long pair_base,
                                                         const T* __restrict__ prec, PairRec R,
                                                         const Cpx<T>* __restrict__ wall,
                                                         int w_stride, int k0, int kc,
                                                         T* __restrict__ D, int st0,
                                                         const long* __restrict__ d_off) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    T* acc = reinterpret_cast<T*>(smem_raw);
    const int ia = blockIdx.x;
    if (ia >= n_atoms) return;
    const int i = atom0 + ia;
    const int s = bd.atom_st[i];
    const int a0 = bd.st_atom_off[s];
    const int nat = bd.st_atom_off[s + 1] - a0;
    const int RD = 6 + 3 * nat;
    const int t1 = bd.types[i];
    const TypeDev& TT = md.t[t1];
    const int ntp = md.n_tp;
    for (int e = threadIdx.x; e < kc * RD; e += blockDim.x) acc[e] = T(0);
    __syncthreads();
    const int k = threadIdx.x;
    if (k < kc) {
        const int fl = pd.fac_local[t1 * pd.K + k0 + k];
        if (fl >= 0) {
            T* a = acc + (size_t)k * RD;
            const Cpx<T>* wi = wall + (size_t)ia * w_stride;
            T s0 = T(0), s1 = T(0), s2 = T(0);
            T st[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
            const int p0 = bd.nb_off[i], p1 = bd.nb_off[i + 1];
            for (int p = p0; p < p1; ++p) {
                const T* rec = prec + (size_t)(p - pair_base) * R.stride;
                if (rec[4] == T(0)) continue;
                const int tp = int(rec[5]);
                const int ga = TT.wz_off[fl * ntp + tp], gb = TT.wz_off[fl * ntp + tp + 1];
                if (ga == gb) continue;
                const T* fn = rec + R.fn();
                const T* fdr = rec + R.fd();
                const T ir = T(1) / rec[3];
                const T ux = rec[0] * ir, uy = rec[1] * ir, uz = rec[2] * ir;
                T q0 = T(0), q1 = T(0), q2 = T(0);
                for (int g = ga; g < gb; ++g) {
                    const int nid = TT.wz_nid[g];
                    const T fv = fn[nid];
                    if (fv < T(1e-20)) continue;
                    const T fdv = fdr[nid];
                    const Cpx<T> W = wi[g];
                    const int key = TT.wz_ylm[g];
                    if (key < 0) {
                        const T r0 = fdv * W.re;
                        q0 += r0 * ux, q1 += r0 * uy, q2 += r0 * uz;
                    } else {
                        const T* y = rec + R.ylm(key);
                        const T ry = (W.re * y[0] - W.im * y[1]) * fdv;
                        q0 += ry * ux + fv * (W.re * y[2] - W.im * y[3]);
                        q1 += ry * uy + fv * (W.re * y[4] - W.im * y[5]);
                        q2 += ry * uz + fv * (W.re * y[6] - W.im * y[7]);
                    }
                }
                s0 += q0, s1 += q1, s2 += q2;
                const int jl = int(rec[6]);
                a[6 + 3 * jl] -= q0;
                a[6 + 3 * jl + 1] -= q1;
                a[6 + 3 * jl + 2] -= q2;
                const T dx = rec[0], dy = rec[1], dz = rec[2];
                st[0] -= q0 * dx, st[1] -= q1 * dy, st[2] -= q2 * dz;
                st[3] -= q0 * dy, st[4] -= q1 * dz, st[5] -= q2 * dx;
            }
            a[6 + 3 * (i - a0)] += s0;
            a[6 + 3 * (i - a0) + 1] += s1;
            a[6 + 3 * (i - a0) + 2] += s2;
            for (int c = 0; c < 6; ++c) a[c] += st[c];
        }
    }
    __syncthreads();
    T* out = D + d_off[s - st0] + ((size_t)(i - a0) * pd.K + k0) * RD;
    for (int e = threadIdx.x; e < kc * RD; e += blockDim.x) out[e] = acc[e];
}

// Gathered monomials of one GEMM group: Ms[s][i][e] = M[i][u_e] over the slot
// entries e, the B operand of the per-factor products.
template <class T>
__global__ void poly_msel_kernel(PolyDev pd, BatchDev bd, int s_first, int n_st, int atom_base,
                                 const T* __restrict__ M, T* __restrict__ Ms, long ms_stride) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    const int s = s_first + blockIdx.z;
    if (e >= pd.E || blockIdx.z >= n_st) return;
    const int a0 = bd.st_atom_off[s];
    if (i >= bd.st_atom_off[s + 1] - a0) return;
    Ms[(size_t)blockIdx.z * ms_stride + (size_t)i * pd.E + e] =
        M[(size_t)(a0 + i - atom_base) * pd.U + pd.ent_u[e]];
}

// X[row][c0 + c] for the structures of one GEMM group: force/stress rows as
// the sum of the column's slot entries of Y[s][r][e], the energy row from the
// factor values; rows weighted, raw energy rows into the scale sums.
template <class T>
__global__ void poly_gather_kernel(PolyDev pd, BatchDev bd, int s_first, int n_st, int atom_base,
                                   const T* __restrict__ Y, long y_stride,
                                   const T* __restrict__ XF, T* __restrict__ X, int ld,
                                   int row_base, const double* __restrict__ roww,
                                   double* __restrict__ xe_sum, double* __restrict__ xe_sq) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = s_first + blockIdx.z;
    if (c >= pd.n_cols || blockIdx.z >= n_st) return;
    const int a0 = bd.st_atom_off[s];
    const int nat = bd.st_atom_off[s + 1] - a0;
    const int RD = 6 + 3 * nat;
    const int rb = bd.st_row_off[s] - row_base;
    const int* ph = pd.col_phi + 3 * (size_t)c;
    const int* pos = pd.col_pos + 3 * (size_t)c;
    const int nf = ph[2] >= 0 ? 3 : 2;
    const T* Ys = Y + (size_t)(s - s_first) * y_stride;
    const int E = pd.E;
    for (int r = blockIdx.y; r < RD; r += gridDim.y) {
        const T* yr = Ys + (size_t)r * E;
        T v = yr[pos[0]] + yr[pos[1]];
        if (nf == 3) v += yr[pos[2]];
        const double w = roww ? roww[rb + 1 + r] : 1.0;
        X[(size_t)(rb + 1 + r) * ld + pd.c0 + c] = T(w * double(v));
    }
    if (blockIdx.y == 0) {
        T e = T(0);
        for (int i = 0; i < nat; ++i) {
            const T* xf = XF + (size_t)(a0 + i - atom_base) * pd.K;
            T p = xf[ph[0]] * xf[ph[1]];
            if (nf == 3) p *= xf[ph[2]];
            e += p;
        }
        const double w = roww ? roww[rb] : 1.0;
        X[(size_t)rb * ld + pd.c0 + c] = T(w * double(e));
        if (xe_sum) {
            atomicAdd(xe_sum + pd.c0 + c, double(e));
            atomicAdd(xe_sq + pd.c0 + c, double(e) * double(e));
        }
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void launch_poly_factor(hipStream_t st, bool fp64, const ModelDev& md, const PolyDev& pd,
                        const BatchDev& bd, int atom0, int n_atoms, const void* xall, int x_stride,
                        void* XF, void* M) {
    const size_t sm = (size_t)pd.K * (fp64 ? 8 : 4);
    if (fp64)
        poly_factor_kernel<double><<<n_atoms, 128, sm, st>>>(
            md, pd, bd, atom0, n_atoms, reinterpret_cast<const double*>(xall), x_stride,
            reinterpret_cast<double*>(XF), reinterpret_cast<double*>(M));
    else
        poly_factor_kernel<float><<<n_atoms, 128, sm, st>>>(
            md, pd, bd, atom0, n_atoms, reinterpret_cast<const float*>(xall), x_stride,
            reinterpret_cast<float*>(XF), reinterpret_cast<float*>(M));
}

void launch_poly_deriv(hipStream_t st, bool fp64, const ModelDev& md, const PolyDev& pd,
                       const BatchDev& bd, int atom0, int n_atoms, long pair_base,
                       const void* prec, const PairRec& R, const void* wall, int w_stride, int k0,
                       int kc, int rd_max, void* D, int st0, const long* d_off) {
    const size_t sm = (size_t)kc * rd_max * (fp64 ? 8 : 4);
    if (fp64) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&poly_deriv_kernel<double>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
        poly_deriv_kernel<double><<<n_atoms, 128, sm, st>>>(
            md, pd, bd, atom0, n_atoms, pair_base, reinterpret_cast<const double*>(prec), R,
            reinterpret_cast<const Cpx<double>*>(wall), w_stride, k0, kc,
            reinterpret_cast<double*>(D), st0, d_off);
    } else {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&poly_deriv_kernel<float>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
        poly_deriv_kernel<float><<<n_atoms, 128, sm, st>>>(
            md, pd, bd, atom0, n_atoms, pair_base, reinterpret_cast<const float*>(prec), R,
            reinterpret_cast<const Cpx<float>*>(wall), w_stride, k0, kc,
            reinterpret_cast<float*>(D), st0, d_off);
    }
}

void launch_poly_msel(hipStream_t st, bool fp64, const PolyDev& pd, const BatchDev& bd,
                      int s_first, int n_st, int atom_base, int n_max, const void* M, void* Ms,
                      long ms_stride) {
    const int B = 256;
    dim3 grid((pd.E + B - 1) / B, n_max, n_st);
    if (fp64)
        poly_msel_kernel<double><<<grid, B, 0, st>>>(pd, bd, s_first, n_st, atom_base,
                                                     reinterpret_cast<const double*>(M),
                                                     reinterpret_cast<double*>(Ms), ms_stride);
    else
        poly_msel_kernel<float><<<grid, B, 0, st>>>(pd, bd, s_first, n_st, atom_base,
                                                    reinterpret_cast<const float*>(M),
                                                    reinterpret_cast<float*>(Ms), ms_stride);
}

void launch_poly_gather(hipStream_t st, bool fp64, const PolyDev& pd, const BatchDev& bd,
                        int s_first, int n_st, int atom_base, const void* Z, long z_stride,
                        const void* XF, void* X, int ld, int row_base, const double* roww,
                        double* xe_sum, double* xe_sq, int rd_max) {
    const int B = 256;
    dim3 grid((pd.n_cols + B - 1) / B, std::min(rd_max, 64), n_st);
    if (fp64)
        poly_gather_kernel<double><<<grid, B, 0, st>>>(
            pd, bd, s_first, n_st, atom_base, reinterpret_cast<const double*>(Z), z_stride,
            reinterpret_cast<const double*>(XF), reinterpret_cast<double*>(X), ld, row_base, roww,
            xe_sum, xe_sq);
    else
        poly_gather_kernel<float><<<grid, B, 0, st>>>(
            pd, bd, s_first, n_st, atom_base, reinterpret_cast<const float*>(Z), z_stride,
            reinterpret_cast<const float*>(XF), reinterpret_cast<float*>(X), ld, row_base, roww,
            xe_sum, xe_sq);
}

}  // namespace plm
