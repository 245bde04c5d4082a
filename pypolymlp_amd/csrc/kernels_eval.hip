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
#include <cstdlib>
#include <cstring>
#include <stdexcept>

#include "device_math.hpp"
#include "device_tables.hpp"
#include "eval_tables.hpp"
#include "kernels.hpp"
#include "pair_data.hpp"

namespace plm {

void check_hip(hipError_t e, const char* what);

// ---------------------------------------------------------------------------
// Per atom: features, products of features, and
//   prod_sum_{e,f}[h] = sum_terms coeff_{e,f} * prod_features * prod_anlm
// One workgroup of NL lanes (one wave, or four for models with many terms)
// per atom; the atom's order parameters and features staged in LDS (the term
// products gather from them at random keys). The terms of all heads are dealt
// to the lanes in equal contiguous chunks (every lane does the same number of
// terms, whatever the heads' lengths), read as one 16-byte key record and one
// coefficient pair per term, one coalesced load per wave and iteration. A
// lane writes the heads that begin and end in its chunk; the heads that cross
// a chunk boundary are summed from the lanes' boundary slots in LDS.
// ---------------------------------------------------------------------------
template <class T, int NL>
__global__ void __launch_bounds__(NL) prodsum_kernel(ModelDev md, EvalDev ed, BatchDev bd,
                                                     const Cpx<T>* __restrict__ a_full,
                                                     int a_stride, Cpx<T>* __restrict__ ps_e,
                                                     Cpx<T>* __restrict__ ps_f, int ps_stride,
                                                     int pair_model) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    // two slots per lane of head sums (4 values each) and their heads, the
    // atom's order parameters (a_stride complex values), then its features
    constexpr int NS = 2 * NL;
    double* sv = reinterpret_cast<double*>(smem_raw);
    int* sh = reinterpret_cast<int*>(sv + 4 * NS);
    Cpx<T>* a = reinterpret_cast<Cpx<T>*>(sh + NS);
    T* xs = reinterpret_cast<T*>(a + a_stride);
    const int i = blockIdx.x;
    if (i >= bd.n_atoms_total) return;
    const int t1 = bd.types[i];
    const TypeDev& TT = md.t[t1];
    const EvalTypeDev& E = ed.t[t1];
    const Cpx<T>* ag = a_full + (size_t)i * a_stride;
    stage_cpx(a, ag, TT.n_nlm);
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
    const int lane = threadIdx.x;
    Cpx<T>* oe = ps_e + (size_t)i * ps_stride;
    Cpx<T>* of = ps_f + (size_t)i * ps_stride;
    const double scale = pair_model ? 0.5 : 1.0;
    // a head that begins and ends inside the lane's chunk is written at once;
    // the chunk's first and last heads (which may continue in the neighbouring
    // chunks) go to the lane's two slots, summed over lanes afterwards
    int cur = -1, first = -1;
    T er = T(0), ei = T(0), fr = T(0), fi = T(0);
    auto to_slot = [&](int sl) {
        sh[sl] = cur;
        sv[4 * sl] = double(er), sv[4 * sl + 1] = double(ei);
        sv[4 * sl + 2] = double(fr), sv[4 * sl + 3] = double(fi);
    };
    sh[2 * lane] = sh[2 * lane + 1] = -1;
    for (int h = lane; h < E.n_heads; h += blockDim.x)
        if (E.pot_off[h + 1] == E.pot_off[h]) oe[h] = {T(0), T(0)}, of[h] = {T(0), T(0)};
    for (int it = 0; it < E.pq_iters; ++it) {
        const int4 k = E.pq_key[(size_t)it * NL + lane];
        const double2 c = E.pq_c[(size_t)it * NL + lane];
        const int head = (unsigned)k.y >> 16;
        if (head == 0xffff) break;  // the rest of this lane's chunk is padding
        if (head != cur) {
            if (cur == first && cur >= 0) {
                to_slot(2 * lane);
            } else if (cur >= 0) {
                oe[cur] = {T(scale * er), T(scale * ei)};
                of[cur] = {T(scale * fr), T(scale * fi)};
            }
            if (first < 0) first = head;
            cur = head, er = ei = fr = fi = T(0);
        }
        const int pd0 = k.x & 0xffff, pd1 = (unsigned)k.x >> 16, pd2 = k.y & 0xffff;
        const int pf0 = k.z & 0xffff, pf1 = (unsigned)k.z >> 16;
        T fval = T(1);
        if (pf0 != 0xffff) fval *= xs[pf0];
        if (pf1 != 0xffff) fval *= xs[pf1];
        Cpx<T> pr = {T(1), T(0)};
        if (pd0 != 0xffff) {
            pr = a[pd0];
            if (pd1 != 0xffff) {
                pr = cmul(pr, a[pd1]);
                if (pd2 != 0xffff) pr = cmul(pr, a[pd2]);
            }
        }
        const T se = fval * T(c.x), sf = fval * T(c.y);
        er += pr.re * se, ei += pr.im * se;
        fr += pr.re * sf, fi += pr.im * sf;
    }
    if (cur >= 0) to_slot(cur == first ? 2 * lane : 2 * lane + 1);
    __syncthreads();
    // slots in term order, empty ones skipped: the first slot of each head sums
    // the slots of that head and writes it
    for (int sl = 2 * lane; sl < 2 * lane + 2; ++sl) {
        const int h = sh[sl];
        if (h < 0) continue;
        int p = sl - 1;
        while (p >= 0 && sh[p] < 0) --p;
        if (p >= 0 && sh[p] == h) continue;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        for (int q = sl; q < NS && (sh[q] == h || sh[q] < 0); ++q) {
            if (sh[q] < 0) continue;
            s0 += sv[4 * q], s1 += sv[4 * q + 1], s2 += sv[4 * q + 2], s3 += sv[4 * q + 3];
        }
        oe[h] = {T(scale * s0), T(scale * s1)};
        of[h] = {T(scale * s2), T(scale * s3)};
    }
}

// ---------------------------------------------------------------------------
// Half-pair loop, record form: one 64-lane workgroup per atom i. The records of up to
// kEvPairs of i's half pairs at a time (radial functions and derivatives,
// Y_lm and its gradient for the reversed vector r_i - r_j - T) are computed
// cooperatively in LDS (pairs_coop, the same arithmetic as the per-lane
// radial / ylm); then 8 lanes per pair split the pair's heads (lane slice s
// takes every 8th head) and the pair's sums are reduced over its 8 lanes.
// Nothing is indexed per lane at run time outside LDS, so no private arrays.
// ---------------------------------------------------------------------------
constexpr int kEvPairs = 8;  // pairs per chunk, 8 lanes each

template <class T>
__global__ void __launch_bounds__(64) pair_rec_kernel(ModelDev md, EvalDev ed, BatchDev bd,
                                                  const int* __restrict__ atom_st,
                                                  const int* __restrict__ hb_off,
                                                  const int* __restrict__ hb_j,
                                                  const double* __restrict__ hb_d,
                                                  const Cpx<T>* __restrict__ ps_e,
                                                  const Cpx<T>* __restrict__ ps_f, int ps_stride,
                                                  double* __restrict__ energy,
                                                  double* __restrict__ forces,
                                                  double* __restrict__ stress, int pair_model) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    T* pairs = reinterpret_cast<T*>(smem_raw);
    const int i = blockIdx.x;
    if (i >= bd.n_atoms_total) return;
    const int p0 = hb_off[i], p1 = hb_off[i + 1];
    if (p0 == p1) return;
    const int t1 = bd.types[i];
    const EvalTypeDev& E = ed.t[t1];
    const PairLayout<T> L(md.n_fn, md.n_lmh);
    BatchDev hb = bd;  // the records are built from the half list
    hb.nb_j = hb_j;
    hb.nb_d = hb_d;
    const int lane = threadIdx.x, pl = lane >> 3, sl = lane & 7;
    const Cpx<T>* Pi_e = ps_e + (size_t)i * ps_stride;
    const Cpx<T>* Pi_f = ps_f + (size_t)i * ps_stride;
    double e_acc = 0.0, fi_acc[3] = {0, 0, 0}, s_acc[6] = {0, 0, 0, 0, 0, 0};
    for (int c0 = p0; c0 < p1; c0 += kEvPairs) {
        const int np = min(kEvPairs, p1 - c0);
        __syncthreads();  // the previous chunk's records are read
        pairs_coop<T>(md, L, pairs, np, c0, hb, t1, 0, true, -1.0, pair_model ? -1 : md.max_l);
        const int k = c0 + pl;
        T* slot = pairs + (size_t)min(pl, np - 1) * L.stride;
        const T* ms = L.misc(slot);
        const bool live = pl < np && ms[4] != T(0);
        // pair model: ea = energy sum, fxa = radial-derivative sum
        T ea = T(0), fxa = T(0), fya = T(0), fza = T(0);
        int j = 0;
        if (live) {
            const int tp = int(ms[5]);
            j = hb_j[k];
            const T* fn = L.fn(slot);
            const T* fd = L.fd(slot);
            const Cpx<T>* Pj_e = ps_e + (size_t)j * ps_stride;
            const Cpx<T>* Pj_f = ps_f + (size_t)j * ps_stride;
            const int h0 = E.hd_off[tp], h1 = E.hd_off[tp + 1];
            if (pair_model) {
                for (int h = h0 + sl; h < h1; h += 8) {
                    const int4 hr = E.hd_rec[h];
                    const int nid = hr.x, hi = hr.z, hj = hr.w & 0x0fffffff;
                    ea += fn[nid] * (Pi_e[hi].re + Pj_e[hj].re);
                    fxa += fd[nid] * (Pi_f[hi].re + Pj_f[hj].re);
                }
            } else {
                const Cpx<T>* Y = L.Y(slot);
                const Cpx<T>* dY = L.dY(slot);
                const T dx = ms[0], dy = ms[1], dz = ms[2];
                const T ir = T(1) / ms[3];
                for (int h = h0 + sl; h < h1; h += 8) {
                    const int4 hr = E.hd_rec[h];
                    const int nid = hr.x;
                    const T fv = fn[nid];
                    if (fv < T(1e-20)) continue;
                    const int key = hr.y;
                    const int hi = hr.z, hj = hr.w & 0x0fffffff;
                    const T sj = (hr.w >> 28) & 1 ? T(-1) : T(1);
                    const T w = (hr.w >> 29) & 1 ? T(0.5) : T(1);
                    const Cpx<T> se = {(Pi_e[hi].re + Pj_e[hj].re * sj) * w,
                                       (Pi_e[hi].im + Pj_e[hj].im * sj) * w};
                    const Cpx<T> sf = {(Pi_f[hi].re + Pj_f[hj].re * sj) * w,
                                       (Pi_f[hi].im + Pj_f[hj].im * sj) * w};
                    const Cpx<T> y = Y[key];
                    ea += fv * (y.re * se.re - y.im * se.im);
                    const T d1r = fd[nid] * y.re * ir, d1i = fd[nid] * y.im * ir;
                    // val_a = -(d1 * d_a + fn * dY_a)
                    const Cpx<T> vx = {-(d1r * dx + fv * dY[3 * key].re), -(d1i * dx + fv * dY[3 * key].im)};
                    const Cpx<T> vy = {-(d1r * dy + fv * dY[3 * key + 1].re), -(d1i * dy + fv * dY[3 * key + 1].im)};
                    const Cpx<T> vz = {-(d1r * dz + fv * dY[3 * key + 2].re), -(d1i * dz + fv * dY[3 * key + 2].im)};
                    fxa += re_mul(vx, sf);
                    fya += re_mul(vy, sf);
                    fza += re_mul(vz, sf);
                }
            }
        }
        // the pair's sums over its 8 lanes
#pragma unroll
        for (int off = 1; off < 8; off <<= 1) {
            ea += __shfl_xor(ea, off);
            fxa += __shfl_xor(fxa, off);
            fya += __shfl_xor(fya, off);
            fza += __shfl_xor(fza, off);
        }
        if (live && sl == 0) {
            // diff_ij: r_i - r_j - T, in FP64 for the forces' directions and the stress
            const double dx = -hb_d[3 * (size_t)k], dy = -hb_d[3 * (size_t)k + 1],
                         dz = -hb_d[3 * (size_t)k + 2];
            const double eij = double(ea);
            double fx, fy, fz;
            if (pair_model) {
                const double fij = -double(fxa) / sqrt(dx * dx + dy * dy + dz * dz);
                fx = fij * dx, fy = fij * dy, fz = fij * dz;
            } else {
                fx = double(fxa), fy = double(fya), fz = double(fza);
            }
            e_acc += eij;
            fi_acc[0] += fx, fi_acc[1] += fy, fi_acc[2] += fz;
            atomicAdd(forces + 3 * (size_t)j, -fx);
            atomicAdd(forces + 3 * (size_t)j + 1, -fy);
            atomicAdd(forces + 3 * (size_t)j + 2, -fz);
            s_acc[0] += dx * fx, s_acc[1] += dy * fy, s_acc[2] += dz * fz;
            s_acc[3] += dx * fy, s_acc[4] += dy * fz, s_acc[5] += dz * fx;
        }
    }
    // wave reduction (64 lanes)
    for (int off = 32; off > 0; off >>= 1) {
        e_acc += __shfl_down(e_acc, off);
        for (int a = 0; a < 3; ++a) fi_acc[a] += __shfl_down(fi_acc[a], off);
        for (int a = 0; a < 6; ++a) s_acc[a] += __shfl_down(s_acc[a], off);
    }
    if (lane == 0) {
        const int s = atom_st[i];
        atomicAdd(energy + s, e_acc);
        for (int a = 0; a < 3; ++a) atomicAdd(forces + 3 * (size_t)i + a, fi_acc[a]);
        for (int a = 0; a < 6; ++a) atomicAdd(stress + 6 * (size_t)s + a, s_acc[a]);
    }
}

// ---------------------------------------------------------------------------
// Half-pair loop, per-lane form: one lane per half pair (i, j) of the whole
// list, the radial functions in registers (NB bounds their number) and Y_lm
// with its gradient generated column by column (LB bounds max l), each key
// used at once on its heads (the heads of a type pair are ordered by key). A
// lane finds its atom i from the half-list offsets; the forces on i and j go
// out as atomics per lane, the energy and stress with one atomic per wave
// when the wave's pairs belong to one structure.
// ---------------------------------------------------------------------------
template <class T, int LB, int NB>
__global__ void __launch_bounds__(256) pair_lane_kernel(ModelDev md, EvalDev ed, BatchDev bd,
                                                   const int* __restrict__ atom_st,
                                                   const int* __restrict__ hb_off,
                                                   const int* __restrict__ hb_j,
                                                   const double* __restrict__ hb_d, int n_half,
                                                   const Cpx<T>* __restrict__ ps_e,
                                                   const Cpx<T>* __restrict__ ps_f, int ps_stride,
                                                   double* __restrict__ energy,
                                                   double* __restrict__ forces,
                                                   double* __restrict__ stress, int pair_model) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double eij = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
    double dx = 0.0, dy = 0.0, dz = 0.0;
    int st = -1;
    if (k < n_half) {
        // atom i of pair k: hb_off[i] <= k < hb_off[i + 1]
        int lo = 0, hi = bd.n_atoms_total;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (hb_off[mid] <= k) lo = mid;
            else hi = mid;
        }
        const int i = lo;
        st = atom_st[i];
        const int t1 = bd.types[i];
        const EvalTypeDev& E = ed.t[t1];
        const Cpx<T>* Pi_e = ps_e + (size_t)i * ps_stride;
        const Cpx<T>* Pi_f = ps_f + (size_t)i * ps_stride;
        const int j = hb_j[k];
        // diff_ij: r_i - r_j - T
        dx = -hb_d[3 * (size_t)k], dy = -hb_d[3 * (size_t)k + 1], dz = -hb_d[3 * (size_t)k + 2];
        const double dis = sqrt(dx * dx + dy * dy + dz * dz);
        if (dis < md.cutoff) {
            const int tp = md.tp_of[t1 * md.n_type + bd.types[j]];
            T fn[NB], fd[NB];
            radial<T>(T(dis), T(md.cutoff), md.rad + (size_t)tp * md.n_fn * 2, md.rad_count[tp], fn, fd);
            const Cpx<T>* Pj_e = ps_e + (size_t)j * ps_stride;
            const Cpx<T>* Pj_f = ps_f + (size_t)j * ps_stride;
            const int h0 = E.hd_off[tp], h1 = E.hd_off[tp + 1];
            if (pair_model) {
                T ef = T(0), ff = T(0);
                for (int h = h0; h < h1; ++h) {
                    const int4 hr = E.hd_rec[h];
                    const int nid = hr.x, hi = hr.z, hj = hr.w & 0x0fffffff;
                    ef += fn[nid] * (Pi_e[hi].re + Pj_e[hj].re);
                    ff += fd[nid] * (Pi_f[hi].re + Pj_f[hj].re);
                }
                eij = double(ef);
                const double fij = -double(ff) / dis;
                fx = fij * dx, fy = fij * dy, fz = fij * dz;
            } else {
                // Y_l^{-m} and its gradient generated column by column, m = max_l
                // down to 0 (column m needs Q(l, m + 1) of the column before), each
                // key used at once on its heads: the recurrences and operation
                // order of ylm(), with no array indexed at run time
                using C = MathConst<T>;
                const T x = T(dx), yy = T(dy), z = T(dz), r = T(dis);
                const T ct = z / r;
                const T rho = dhypot(x, yy);
                T cp = T(1), sp = T(0);
                if (rho > T(0)) cp = x / rho, sp = yy / rho;
                const T st = dsqrt(T(1) - ct * ct);
                const T invr = T(1) / r;
                const T ir = T(1) / T(dis);
                T cm[LB + 1], sm[LB + 1];  // cos, sin (m phi) by the Chebyshev recurrence
                cm[0] = T(1), sm[0] = T(0);
                {
                    T c1 = T(1), c2 = cp, s1 = T(0), s2 = -sp;
                    const T tc = T(2) * c2;
#pragma unroll
                    for (int mp = 1; mp <= LB; ++mp) {
                        const T sv = tc * s1 - s2;
                        const T cv = tc * c1 - c2;
                        c2 = c1, c1 = cv, s2 = s1, s1 = sv;
                        cm[mp] = cv, sm[mp] = sv;
                    }
                }
                const int lmax = md.max_l;
                const int* ko = E.hd_koff + (size_t)tp * (md.n_lmh + 1);
                T ea = T(0), fxa = T(0), fya = T(0), fza = T(0);
                auto heads = [&](int key, Cpx<T> y, Cpx<T> gx, Cpx<T> gy, Cpx<T> gz) {
                    for (int h = ko[key]; h < ko[key + 1]; ++h) {
                        const int4 hr = E.hd_rec[h];
                        const int nid = hr.x;
                        const T fv = fn[nid];
                        if (fv < T(1e-20)) continue;
                        const int hi = hr.z, hj = hr.w & 0x0fffffff;
                        const T sj = (hr.w >> 28) & 1 ? T(-1) : T(1);
                        const T w = (hr.w >> 29) & 1 ? T(0.5) : T(1);
                        const Cpx<T> se = {(Pi_e[hi].re + Pj_e[hj].re * sj) * w,
                                           (Pi_e[hi].im + Pj_e[hj].im * sj) * w};
                        const Cpx<T> sf = {(Pi_f[hi].re + Pj_f[hj].re * sj) * w,
                                           (Pi_f[hi].im + Pj_f[hj].im * sj) * w};
                        ea += fv * (y.re * se.re - y.im * se.im);
                        const T d1r = fd[nid] * y.re * ir, d1i = fd[nid] * y.im * ir;
                        // val_a = -(d1 * d_a + fn * dY_a)
                        const Cpx<T> vx = {-(d1r * x + fv * gx.re), -(d1i * x + fv * gx.im)};
                        const Cpx<T> vy = {-(d1r * yy + fv * gy.re), -(d1i * yy + fv * gy.im)};
                        const Cpx<T> vz = {-(d1r * z + fv * gz.re), -(d1i * z + fv * gz.im)};
                        fxa += re_mul(vx, sf);
                        fya += re_mul(vy, sf);
                        fza += re_mul(vz, sf);
                    }
                };
                T qn[LB + 1];  // Q(l, m + 1)
#pragma unroll
                for (int l = 0; l <= LB; ++l) qn[l] = T(0);
#pragma unroll
                for (int m = LB; m >= 0; --m) {
                    if (m > lmax) continue;
                    // normalised associated Legendre column m: P and Q = P / sin(theta)
                    T pc[LB + 1], qc[LB + 1];
                    T pmm = C::sqrt_inv_2pi, qmm = T(0);
                    if (m >= 1) {
                        pmm = -st * C::sqrt3div2 * C::sqrt_inv_2pi;
                        qmm = -C::sqrt3div2 * C::sqrt_inv_2pi;
#pragma unroll
                        for (int l = 2; l <= m; ++l) {
                            const T c1 = -dsqrt(T(1) + T(0.5) / T(l)) * st;
                            pmm = c1 * pmm;
                            qmm = c1 * qmm;
                        }
                    }
#pragma unroll
                    for (int l = 0; l <= LB; ++l) pc[l] = T(0), qc[l] = T(0);
                    pc[m] = pmm, qc[m] = qmm;
                    if (m + 1 <= LB && m + 1 <= lmax) {
                        if (m == 0) {
                            pc[1] = ct * C::sqrt3 * C::sqrt_inv_2pi;
                            qc[1] = T(0);
                        } else {
                            const T c2 = dsqrt(T(2) * (T(m + 1) - T(1)) + T(3)) * ct;
                            pc[m + 1] = c2 * pmm;
                            qc[m + 1] = c2 * qmm;
                        }
#pragma unroll
                        for (int l = m + 2; l <= LB; ++l) {
                            if (l > lmax) break;
                            const T ls = T(l * l), lm1s = T((l - 1) * (l - 1)), msq = T(m * m);
                            const T alm = dsqrt((T(4) * ls - T(1)) / (ls - msq));
                            const T blm = -dsqrt((lm1s - msq) / (T(4) * lm1s - T(1)));
                            pc[l] = alm * (ct * pc[l - 1] + blm * pc[l - 2]);
                            qc[l] = alm * (ct * qc[l - 1] + blm * qc[l - 2]);
                        }
                    }
#pragma unroll
                    for (int l = m; l <= LB; ++l) {
                        if (l > lmax) break;
                        if (m == 0) {
                            const T common =
                                (l >= 1) ? qn[l] * st * invr * dsqrt(T(0.5) * T(l) * T(l + 1)) : T(0);
                            heads(lm2i(l, 0) + l, Cpx<T>{pc[l] * C::half_sqrt2, T(0)},
                                  Cpx<T>{common * ct * cp, T(0)}, Cpx<T>{common * ct * sp, T(0)},
                                  Cpx<T>{-common * st, T(0)});
                        } else {
                            const T c = cm[m], sv = sm[m];
                            const T sign = (m % 2 == 1) ? T(-1) : T(1);
                            const T tmp = sign * pc[l] * C::half_sqrt2;
                            const T cre = c * C::half_sqrt2 * invr, cim = sv * C::half_sqrt2 * invr;
                            T dth = T(m) * ct * qc[l];
                            if (m != l) dth += dsqrt(T((l - m) * (l + m + 1))) * qn[l] * st;
                            const T dph = T(m) * qc[l];
                            const T ure = dth * ct * cp, uim = -dph * sp;
                            const T vre = dth * ct * sp, vim = dph * cp;
                            const T wre = -dth * st;
                            heads(lm2i(l, -m) + l, Cpx<T>{tmp * c, -tmp * sv},
                                  Cpx<T>{sign * (cre * ure - cim * uim), -sign * (cre * uim + cim * ure)},
                                  Cpx<T>{sign * (cre * vre - cim * vim), -sign * (cre * vim + cim * vre)},
                                  Cpx<T>{sign * (cre * wre), -sign * (cim * wre)});
                        }
                    }
#pragma unroll
                    for (int l = 0; l <= LB; ++l) qn[l] = qc[l];
                }
                eij = double(ea), fx = double(fxa), fy = double(fya), fz = double(fza);
            }
            atomicAdd(forces + 3 * (size_t)i, fx);
            atomicAdd(forces + 3 * (size_t)i + 1, fy);
            atomicAdd(forces + 3 * (size_t)i + 2, fz);
            atomicAdd(forces + 3 * (size_t)j, -fx);
            atomicAdd(forces + 3 * (size_t)j + 1, -fy);
            atomicAdd(forces + 3 * (size_t)j + 2, -fz);
        }
    }
    double sv[7] = {eij, dx * fx, dy * fy, dz * fz, dx * fy, dy * fz, dz * fx};
    const int s0 = __builtin_amdgcn_readfirstlane(st);
    if (s0 < 0) return;  // lane 0 past the end: so is the whole wave
    if (__all(st == s0 || st < 0)) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
#pragma unroll
            for (int a = 0; a < 7; ++a) sv[a] += __shfl_down(sv[a], off);
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(energy + s0, sv[0]);
            for (int a = 0; a < 6; ++a) atomicAdd(stress + 6 * (size_t)s0 + a, sv[1 + a]);
        }
    } else if (st >= 0) {
        atomicAdd(energy + st, sv[0]);
        for (int a = 0; a < 6; ++a) atomicAdd(stress + 6 * (size_t)st + a, sv[1 + a]);
    }
}

template <class T, int NL>
void launch_prodsum_t(hipStream_t st, const ModelDev& md, const EvalDev& ed, const BatchDev& bd,
                      const void* a_full, int a_stride, void* ps_e, void* ps_f, int ps_stride,
                      int max_lin, int pair_model) {
    const size_t sm = 2 * NL * (4 * sizeof(double) + sizeof(int)) +
                      (size_t)a_stride * 2 * sizeof(T) + (size_t)std::max(1, max_lin) * sizeof(T);
    if (sm > 160 * 1024) throw std::invalid_argument("per-atom evaluation data exceed LDS");
    check_hip(hipFuncSetAttribute(reinterpret_cast<const void*>(&prodsum_kernel<T, NL>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm),
              "prodsum shared memory");
    prodsum_kernel<T, NL><<<bd.n_atoms_total, NL, sm, st>>>(
        md, ed, bd, reinterpret_cast<const Cpx<T>*>(a_full), a_stride,
        reinterpret_cast<Cpx<T>*>(ps_e), reinterpret_cast<Cpx<T>*>(ps_f), ps_stride, pair_model);
}

template <class T>
void launch_prodsum_lanes(hipStream_t st, const ModelDev& md, const EvalDev& ed,
                          const BatchDev& bd, const void* a_full, int a_stride, void* ps_e,
                          void* ps_f, int ps_stride, int max_lin, int pair_model) {
    // the term tables were laid out for ed.pq_lanes lanes
    if (ed.pq_lanes == 256)
        launch_prodsum_t<T, 256>(st, md, ed, bd, a_full, a_stride, ps_e, ps_f, ps_stride, max_lin,
                                 pair_model);
    else if (ed.pq_lanes == 64)
        launch_prodsum_t<T, 64>(st, md, ed, bd, a_full, a_stride, ps_e, ps_f, ps_stride, max_lin,
                                pair_model);
    else
        throw std::invalid_argument("contraction lanes must be 64 or 256");
}

void launch_prodsum(hipStream_t st, bool fp64, const ModelDev& md, const EvalDev& ed,
                    const BatchDev& bd, const void* a_full, int a_stride, void* ps_e, void* ps_f,
                    int ps_stride, int max_lin, int pair_model) {
    if (fp64)
        launch_prodsum_lanes<double>(st, md, ed, bd, a_full, a_stride, ps_e, ps_f, ps_stride,
                                     max_lin, pair_model);
    else
        launch_prodsum_lanes<float>(st, md, ed, bd, a_full, a_stride, ps_e, ps_f, ps_stride,
                                    max_lin, pair_model);
}

template <class T>
void launch_pair_rec_t(hipStream_t st, const ModelDev& md, const EvalDev& ed, const BatchDev& bd,
                   const int* atom_st, const int* hb_off, const int* hb_j, const double* hb_d,
                   const void* ps_e, const void* ps_f, int ps_stride, double* energy,
                   double* forces, double* stress, int pair_model) {
    const size_t sm = (size_t)kEvPairs * PairLayout<T>(md.n_fn, md.n_lmh).stride * sizeof(T);
    if (sm > 160 * 1024) throw std::invalid_argument("evaluation pair records exceed LDS");
    check_hip(hipFuncSetAttribute(reinterpret_cast<const void*>(&pair_rec_kernel<T>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm),
              "pair shared memory");
    pair_rec_kernel<T><<<bd.n_atoms_total, 64, sm, st>>>(
        md, ed, bd, atom_st, hb_off, hb_j, hb_d, reinterpret_cast<const Cpx<T>*>(ps_e),
        reinterpret_cast<const Cpx<T>*>(ps_f), ps_stride, energy, forces, stress, pair_model);
}

template <class T, int LB, int NB>
void launch_pair_lb(hipStream_t st, const ModelDev& md, const EvalDev& ed, const BatchDev& bd,
                    const int* atom_st, const int* hb_off, const int* hb_j, const double* hb_d,
                    int n_half, const void* ps_e, const void* ps_f, int ps_stride, double* energy,
                    double* forces, double* stress, int pair_model) {
    if (n_half <= 0) return;
    pair_lane_kernel<T, LB, NB><<<(n_half + 255) / 256, 256, 0, st>>>(
        md, ed, bd, atom_st, hb_off, hb_j, hb_d, n_half, reinterpret_cast<const Cpx<T>*>(ps_e),
        reinterpret_cast<const Cpx<T>*>(ps_f), ps_stride, energy, forces, stress, pair_model);
}

template <class T>
void launch_pair_lane_t(hipStream_t st, const ModelDev& md, const EvalDev& ed,
                        const BatchDev& bd, const int* atom_st, const int* hb_off,
                        const int* hb_j, const double* hb_d, int n_half, const void* ps_e,
                        const void* ps_f, int ps_stride, double* energy, double* forces,
                        double* stress, int pair_model) {
    const int L = md.max_l;
    // radial arrays of 16 entries where the model has at most 16 radial
    // functions and max_l <= 4: in FP64 they stay in registers (32 spill 528
    // bytes per lane to scratch; 2.26 against 2.53 ms per 32,000 atoms), in
    // FP32 the lower register count holds 4 waves per SIMD instead of 3 (1.60
    // against 1.66 ms); at max_l 8 the 16-entry form's registers cost more than
    // the spill (4.40 against 3.90 ms), so it keeps 32
    if (L <= 4 && md.n_fn <= 16) {
        launch_pair_lb<T, 4, 16>(st, md, ed, bd, atom_st, hb_off, hb_j, hb_d, n_half, ps_e, ps_f,
                                 ps_stride, energy, forces, stress, pair_model);
        return;
    }
    if (md.n_fn > 32)
        throw std::invalid_argument("the per-lane pair loop supports at most 32 radial functions");
    if (L <= 4)
        launch_pair_lb<T, 4, 32>(st, md, ed, bd, atom_st, hb_off, hb_j, hb_d, n_half, ps_e, ps_f, ps_stride,
                                 energy, forces, stress, pair_model);
    else if (L <= 8)
        launch_pair_lb<T, 8, 32>(st, md, ed, bd, atom_st, hb_off, hb_j, hb_d, n_half, ps_e, ps_f, ps_stride,
                                 energy, forces, stress, pair_model);
    else if (L <= 12)
        launch_pair_lb<T, 12, 32>(st, md, ed, bd, atom_st, hb_off, hb_j, hb_d, n_half, ps_e, ps_f, ps_stride,
                                  energy, forces, stress, pair_model);
    else
        throw std::invalid_argument("the per-lane pair loop supports max_l <= 12");
}

// The per-lane form where it applies (max_l <= 12, at most 32 radial
// functions), the record form beyond; PLM_EVAL_PAIR=lane / records forces one.
void launch_pair(hipStream_t st, bool fp64, const ModelDev& md, const EvalDev& ed,
                 const BatchDev& bd, const int* atom_st, const int* hb_off, const int* hb_j,
                 const double* hb_d, int n_half, const void* ps_e, const void* ps_f, int ps_stride,
                 double* energy, double* forces, double* stress, int pair_model) {
    if (n_half <= 0) return;
    bool lane = md.max_l <= 12 && md.n_fn <= 32;
    if (const char* e = getenv("PLM_EVAL_PAIR")) {
        if (!strcmp(e, "lane")) lane = true;
        else if (!strcmp(e, "records")) lane = false;
        else throw std::invalid_argument("PLM_EVAL_PAIR: lane or records");
    }
    if (lane && fp64)
        launch_pair_lane_t<double>(st, md, ed, bd, atom_st, hb_off, hb_j, hb_d, n_half, ps_e, ps_f,
                                   ps_stride, energy, forces, stress, pair_model);
    else if (lane)
        launch_pair_lane_t<float>(st, md, ed, bd, atom_st, hb_off, hb_j, hb_d, n_half, ps_e, ps_f,
                                  ps_stride, energy, forces, stress, pair_model);
    else if (fp64)
        launch_pair_rec_t<double>(st, md, ed, bd, atom_st, hb_off, hb_j, hb_d, ps_e, ps_f,
                                  ps_stride, energy, forces, stress, pair_model);
    else
        launch_pair_rec_t<float>(st, md, ed, bd, atom_st, hb_off, hb_j, hb_d, ps_e, ps_f,
                                 ps_stride, energy, forces, stress, pair_model);
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
