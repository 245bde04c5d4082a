// Polynomial columns of X (products of 2 or 3 linear features, the
// "factors"), first half: factor values, monomials and factor derivatives.
//
// For a column c = prod_s x_{f_s} (Model::model_order2/3, model.cpp:207-267):
//   force/stress rows:  X[r, c] = sum_i sum_s D_{f_s}[r, i] * m_{c\s}(i)
//   energy row:         X[e, c] = sum_i prod_s x_{f_s}(i)
// where D_f[r, i] = d x_f(i) / d xi_r is the derivative of factor f at atom i
// (rows r: 6 stress, then 3N forces) and m_{c\s}(i) the product of the other
// factors at atom i. Each (slot s, column) pair is an entry e = (f, u) with
// f the slot's factor and u the monomial of the others; the products of the
// entries and the column sums are in kernels_polyz.hip.
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

// D[i][f][r] for the factors f in [k0, k0 + kc) (kc <= 128): one 256-thread
// workgroup per atom, blockDim / kc threads per factor, each walking every
// (blockDim / kc)-th pair of the atom (rows of neighbours j added with LDS
// atomics, the rows of atom i and the 6 stress rows summed in registers and
// added once), then a coalesced write of the atom's rows.
template <class T>
__global__ void __launch_bounds__(256) poly_deriv_kernel(ModelDev md, PolyDev pd, BatchDev bd,
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
    // thread -> (factor k, pair slice sl): the S slices of a factor take every
    // S-th pair of the atom and add into the factor's rows with LDS atomics
    const int S = blockDim.x / kc;
    const int k = threadIdx.x % kc, sl = threadIdx.x / kc;
    if (sl < S) {
        const int fl = pd.fac_local[t1 * pd.K + k0 + k];
        if (fl >= 0) {
            T* a = acc + (size_t)k * RD;
            const Cpx<T>* wi = wall + (size_t)ia * w_stride;
            T s0 = T(0), s1 = T(0), s2 = T(0);
            T st[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
            const int p0 = bd.nb_off[i], p1 = bd.nb_off[i + 1];
            // record headers (d, r | valid, tp, j) as two 4-element loads; the
            // next pair's header loads while this pair is processed
            using V4 = typename Vec4Of<T>::type;
            V4 h0n{}, h1n{};
            if (p0 + sl < p1) {
                const V4* hr = reinterpret_cast<const V4*>(prec + (size_t)(p0 + sl - pair_base) * R.stride);
                h0n = hr[0], h1n = hr[1];
            }
            for (int p = p0 + sl; p < p1; p += S) {
                const T* rec = prec + (size_t)(p - pair_base) * R.stride;
                const V4 h0 = h0n, h1 = h1n;
                if (p + S < p1) {
                    const V4* hr = reinterpret_cast<const V4*>(rec + (size_t)S * R.stride);
                    h0n = hr[0], h1n = hr[1];
                }
                if (h1.x == T(0)) continue;
                const int tp = int(h1.y);
                const int ga = TT.wz_off[fl * ntp + tp], gb = TT.wz_off[fl * ntp + tp + 1];
                if (ga == gb) continue;
                const T* fn = rec + R.fn();
                const T* fdr = rec + R.fd();
                const T ir = T(1) / h0.w;
                const T ux = h0.x * ir, uy = h0.y * ir, uz = h0.z * ir;
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
                const int jl = int(h1.z);
                atomicAdd(a + 6 + 3 * jl, -q0);
                atomicAdd(a + 6 + 3 * jl + 1, -q1);
                atomicAdd(a + 6 + 3 * jl + 2, -q2);
                const T dx = h0.x, dy = h0.y, dz = h0.z;
                st[0] -= q0 * dx, st[1] -= q1 * dy, st[2] -= q2 * dz;
                st[3] -= q0 * dy, st[4] -= q1 * dz, st[5] -= q2 * dx;
            }
            atomicAdd(a + 6 + 3 * (i - a0), s0);
            atomicAdd(a + 6 + 3 * (i - a0) + 1, s1);
            atomicAdd(a + 6 + 3 * (i - a0) + 2, s2);
            for (int c = 0; c < 6; ++c) atomicAdd(a + c, st[c]);
        }
    }
    __syncthreads();
    T* out = D + d_off[s - st0] + ((size_t)(i - a0) * pd.K + k0) * RD;
    for (int e = threadIdx.x; e < kc * RD; e += blockDim.x) out[e] = acc[e];
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
        poly_deriv_kernel<double><<<n_atoms, 256, sm, st>>>(
            md, pd, bd, atom0, n_atoms, pair_base, reinterpret_cast<const double*>(prec), R,
            reinterpret_cast<const Cpx<double>*>(wall), w_stride, k0, kc,
            reinterpret_cast<double*>(D), st0, d_off);
    } else {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&poly_deriv_kernel<float>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
        poly_deriv_kernel<float><<<n_atoms, 256, sm, st>>>(
            md, pd, bd, atom0, n_atoms, pair_base, reinterpret_cast<const float*>(prec), R,
            reinterpret_cast<const Cpx<float>*>(wall), w_stride, k0, kc,
            reinterpret_cast<float*>(D), st0, d_off);
    }
}

}  // namespace plm
