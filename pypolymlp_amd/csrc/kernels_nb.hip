// Cell-list neighbour search for large cells with the reference's pair set.
//
// Atoms are binned on their wrapped fractional coordinates. For atom i and
// each of the 27 neighbouring bins, an image of atom j is the integer lattice
// shift k = s - w_j + w_i (s: bin wrap, w: floor of the fractional
// coordinates). The pair is kept only if k is one of the reference's lattice
// translations (NeighborCell::find_trans, neighbor_cell.cpp:236-255) and
// passes the reference's distance predicate evaluated with the same rounded
// operations as NeighborFull / NeighborHalf (neighbor_full.cpp:35-48,
// neighbor_half.cpp:35-76). Each atom's list is then sorted by (j, translation
// index), which is the reference's order, so the lists are bit-identical.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace plm {

struct BinGrid {
    int nb[3];
    int mexp[3];          // translation index box: k_a in [-mexp_a, mexp_a]
    double inv[9];        // inverse of the axis used for translations (row-major)
    const int* tindex;    // [(2m0+1)(2m1+1)(2m2+1)] -> translation index or -1
};

__global__ void bin_assign_kernel(int n, const double* __restrict__ pos, BinGrid g,
                                  int* __restrict__ bin, int* __restrict__ wrap,
                                  int* __restrict__ bin_count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* r = pos + 3 * (size_t)i;
    int b[3];
    for (int a = 0; a < 3; ++a) {
        const double f = g.inv[3 * a] * r[0] + g.inv[3 * a + 1] * r[1] + g.inv[3 * a + 2] * r[2];
        const double w = floor(f);
        const double fw = f - w;
        int bb = (int)(fw * g.nb[a]);
        bb = min(max(bb, 0), g.nb[a] - 1);
        b[a] = bb;
        wrap[3 * (size_t)i + a] = (int)w;
    }
    const int id = (b[0] * g.nb[1] + b[1]) * g.nb[2] + b[2];
    bin[i] = id;
    atomicAdd(bin_count + id, 1);
}

__global__ void bin_scatter_kernel(int n, const int* __restrict__ bin, const int* __restrict__ start,
                                   int* __restrict__ cursor, int* __restrict__ sorted) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b = bin[i];
    const int slot = atomicAdd(cursor + b, 1);
    sorted[start[b] + slot] = i;
}

__device__ __forceinline__ bool pred(const double* pi, const double* pj, const double* tr,
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

// One thread per atom. Count pass (nb_off == nullptr) or fill + sort pass.
__global__ void bin_pairs_kernel(int n, int half, const double* __restrict__ pos,
                                 const int* __restrict__ bin, const int* __restrict__ wrap,
                                 const int* __restrict__ start, const int* __restrict__ sorted,
                                 BinGrid g, const double* __restrict__ trans, double cut2,
                                 double tol2, double tol, int* __restrict__ count,
                                 const int* __restrict__ nb_off, int* __restrict__ nb_j,
                                 int* __restrict__ nb_t, double* __restrict__ nb_d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* pi = pos + 3 * (size_t)i;
    const int bi = bin[i];
    const int b0 = bi / (g.nb[1] * g.nb[2]), b1 = (bi / g.nb[2]) % g.nb[1], b2 = bi % g.nb[2];
    const int wi0 = wrap[3 * (size_t)i], wi1 = wrap[3 * (size_t)i + 1], wi2 = wrap[3 * (size_t)i + 2];
    const int M0 = 2 * g.mexp[0] + 1, M1 = 2 * g.mexp[1] + 1, M2 = 2 * g.mexp[2] + 1;
    int c = 0;
    const int out0 = nb_off ? nb_off[i] : 0;
    int out = out0;
    for (int d0 = -1; d0 <= 1; ++d0) {
        const int q0 = b0 + d0, s0 = (q0 < 0) ? -1 : (q0 >= g.nb[0] ? 1 : 0);
        const int c0 = q0 - s0 * g.nb[0];
        for (int d1 = -1; d1 <= 1; ++d1) {
            const int q1 = b1 + d1, s1 = (q1 < 0) ? -1 : (q1 >= g.nb[1] ? 1 : 0);
            const int c1 = q1 - s1 * g.nb[1];
            for (int d2 = -1; d2 <= 1; ++d2) {
                const int q2 = b2 + d2, s2 = (q2 < 0) ? -1 : (q2 >= g.nb[2] ? 1 : 0);
                const int c2 = q2 - s2 * g.nb[2];
                const int cb = (c0 * g.nb[1] + c1) * g.nb[2] + c2;
                for (int e = start[cb]; e < start[cb + 1]; ++e) {
                    const int j = sorted[e];
                    if (half && j > i) continue;
                    const int k0 = s0 - wrap[3 * (size_t)j] + wi0;
                    const int k1 = s1 - wrap[3 * (size_t)j + 1] + wi1;
                    const int k2 = s2 - wrap[3 * (size_t)j + 2] + wi2;
                    if (k0 < -g.mexp[0] || k0 > g.mexp[0] || k1 < -g.mexp[1] || k1 > g.mexp[1] ||
                        k2 < -g.mexp[2] || k2 > g.mexp[2])
                        continue;
                    const int t = g.tindex[((k0 + g.mexp[0]) * M1 + (k1 + g.mexp[1])) * M2 +
                                           (k2 + g.mexp[2])];
                    if (t < 0) continue;
                    const double* tr = trans + 3 * (size_t)t;
                    double d[3];
                    if (j == i) {
                        if (!half) {
                            if (!pred(pi, pi, tr, cut2, tol2, d)) continue;
                        } else {
                            const double dx = tr[0], dy = tr[1], dz = tr[2];
                            const double r2 = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)),
                                                        __dmul_rn(dz, dz));
                            if (!(r2 < cut2 && r2 > tol2)) continue;
                            bool keep;
                            if (dz >= tol) keep = true;
                            else if (fabs(dz) < tol && dy >= tol) keep = true;
                            else if (fabs(dz) < tol && fabs(dy) < tol && dx >= tol) keep = true;
                            else keep = false;
                            if (!keep) continue;
                            d[0] = dx, d[1] = dy, d[2] = dz;
                        }
                    } else {
                        if (!pred(pi, pos + 3 * (size_t)j, tr, cut2, tol2, d)) continue;
                    }
                    if (nb_off) {
                        nb_j[out] = j;
                        nb_t[out] = t;
                        nb_d[3 * (size_t)out] = d[0];
                        nb_d[3 * (size_t)out + 1] = d[1];
                        nb_d[3 * (size_t)out + 2] = d[2];
                        ++out;
                    }
                    ++c;
                }
            }
        }
    }
    (void)M0;
    if (!nb_off) {
        count[i] = c;
        return;
    }
    // insertion sort of this atom's segment by (j, t); in the half list the
    // self images (j == i) come after all j < i, which the key order gives too
    for (int a = out0 + 1; a < out; ++a) {
        const int kj = nb_j[a], kt = nb_t[a];
        const double e0 = nb_d[3 * (size_t)a], e1 = nb_d[3 * (size_t)a + 1], e2 = nb_d[3 * (size_t)a + 2];
        int b = a - 1;
        while (b >= out0 && (nb_j[b] > kj || (nb_j[b] == kj && nb_t[b] > kt))) {
            nb_j[b + 1] = nb_j[b];
            nb_t[b + 1] = nb_t[b];
            nb_d[3 * (size_t)(b + 1)] = nb_d[3 * (size_t)b];
            nb_d[3 * (size_t)(b + 1) + 1] = nb_d[3 * (size_t)b + 1];
            nb_d[3 * (size_t)(b + 1) + 2] = nb_d[3 * (size_t)b + 2];
            --b;
        }
        nb_j[b + 1] = kj;
        nb_t[b + 1] = kt;
        nb_d[3 * (size_t)(b + 1)] = e0;
        nb_d[3 * (size_t)(b + 1) + 1] = e1;
        nb_d[3 * (size_t)(b + 1) + 2] = e2;
    }
}

void launch_bin_assign(hipStream_t st, int n, const double* pos, const BinGridHost& gh, int* bin,
                       int* wrap, int* bin_count) {
    BinGrid g;
    for (int a = 0; a < 3; ++a) g.nb[a] = gh.nb[a], g.mexp[a] = gh.mexp[a];
    for (int a = 0; a < 9; ++a) g.inv[a] = gh.inv[a];
    g.tindex = gh.tindex;
    bin_assign_kernel<<<(n + 255) / 256, 256, 0, st>>>(n, pos, g, bin, wrap, bin_count);
}

void launch_bin_scatter(hipStream_t st, int n, const int* bin, const int* start, int* cursor,
                        int* sorted) {
    bin_scatter_kernel<<<(n + 255) / 256, 256, 0, st>>>(n, bin, start, cursor, sorted);
}

void launch_bin_pairs(hipStream_t st, int n, int half, const double* pos, const int* bin,
                      const int* wrap, const int* start, const int* sorted, const BinGridHost& gh,
                      const double* trans, double cut2, double tol2, double tol, int* count,
                      const int* nb_off, int* nb_j, int* nb_t, double* nb_d) {
    BinGrid g;
    for (int a = 0; a < 3; ++a) g.nb[a] = gh.nb[a], g.mexp[a] = gh.mexp[a];
    for (int a = 0; a < 9; ++a) g.inv[a] = gh.inv[a];
    g.tindex = gh.tindex;
    bin_pairs_kernel<<<(n + 127) / 128, 128, 0, st>>>(n, half, pos, bin, wrap, start, sorted, g,
                                                      trans, cut2, tol2, tol, count, nb_off, nb_j,
                                                      nb_t, nb_d);
}

}  // namespace plm
