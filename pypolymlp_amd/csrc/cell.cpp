// Lattice translations and position rewrapping for periodic neighbour lists.
//
// Restates NeighborCell (src/cxx/src/compute/neighbor_cell.cpp:11-257) in FP64
// with the same operation order, so that the translation vectors, the
// rewrapped positions and therefore the neighbour sets are bit-identical:
//  * Gram entries of the lattice vectors (columns of axis) and the "needs
//    reduction" test |int(a_i . a_j)| > |a_k|^2 / 2 (compute_metric, :24-44;
//    the integer truncation is the reference's behaviour, see iabs below);
//  * the 1-D axis reduction loop (refine_axis, :126-168, at most 100 sweeps);
//  * only when the cell was reduced: inverse axis, fractional coordinates,
//    p -= floor(p), back to Cartesian (find_trans, :203-235);
//  * translations T = A (i, j, k) with |i| <= ceil(rc / |a_i|) + 1 kept when
//    |T| < max vertex distance of the cell + rc (:236-255).
// This file is compiled with -ffp-contract=off.
#include "cell.hpp"

#include <cmath>

namespace plm {

namespace {

inline int iabs(double x) { return std::abs(static_cast<int>(x)); }

struct Cell {
    double ax[3][3];  // ax[i][col]
    double d00, d11, d22, d01, d02, d12;
    bool r01, r02, r12;
    int rsum;

    void metric() {
        double a0[3], a1[3], a2[3];
        for (int i = 0; i < 3; ++i) a0[i] = ax[i][0], a1[i] = ax[i][1], a2[i] = ax[i][2];
        auto dot = [](const double* u, const double* v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; };
        d00 = dot(a0, a0), d11 = dot(a1, a1), d22 = dot(a2, a2);
        d01 = dot(a0, a1), d02 = dot(a0, a2), d12 = dot(a1, a2);
        // The reference's unqualified abs() on these doubles resolves to the C
        // library's int abs(int): the Gram entries are truncated toward zero
        // before the comparison. Reproduced for identical neighbour sets.
        r01 = (iabs(d01) > 0.5 * d00 || iabs(d01) > 0.5 * d11);
        r02 = (iabs(d02) > 0.5 * d00 || iabs(d02) > 0.5 * d22);
        r12 = (iabs(d12) > 0.5 * d11 || iabs(d12) > 0.5 * d22);
        rsum = int(r01) + int(r02) + int(r12);
    }
    template <class V>
    void cart(V i, V j, V k, double* out) const {
        out[0] = ax[0][0] * i + ax[0][1] * j + ax[0][2] * k;
        out[1] = ax[1][0] * i + ax[1][1] * j + ax[1][2] * k;
        out[2] = ax[2][0] * i + ax[2][1] * j + ax[2][2] * k;
    }
    template <class V>
    double dist(V i, V j, V k) const {
        double v[3];
        cart(i, j, k, v);
        return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    void replace(int i, int j, int k, int col) {
        double v[3];
        cart(i, j, k, v);
        for (int a = 0; a < 3; ++a) ax[a][col] = v[a];
        metric();
    }
    // optimize_1d_axis{0,1,2}: minimise along one axis with the others fixed
    void opt_axis(int axis_id, int v1, int v2, int out[3]) const {
        double imin;
        if (axis_id == 0) imin = -(d01 * v1 + d02 * v2) / d00;
        else if (axis_id == 1) imin = -(d01 * v1 + d12 * v2) / d11;
        else imin = -(d02 * v1 + d12 * v2) / d22;
        const int i = (int)std::round(imin);
        if (axis_id == 0) out[0] = i, out[1] = v1, out[2] = v2;
        else if (axis_id == 1) out[0] = v1, out[1] = i, out[2] = v2;
        else out[0] = v1, out[1] = v2, out[2] = i;
    }
    void refine() {
        int v[3];
        int iter = 0;
        while (rsum > 0 && iter < 100) {
            if (r01) {
                if (d00 > d11) opt_axis(1, 1, 0, v), replace(v[0], v[1], v[2], 0);
                else opt_axis(0, 1, 0, v), replace(v[0], v[1], v[2], 1);
            }
            if (r02) {
                if (d00 > d22) opt_axis(2, 1, 0, v), replace(v[0], v[1], v[2], 0);
                else opt_axis(0, 0, 1, v), replace(v[0], v[1], v[2], 2);
            }
            if (r12) {
                if (d11 > d22) opt_axis(2, 0, 1, v), replace(v[0], v[1], v[2], 1);
                else opt_axis(1, 0, 1, v), replace(v[0], v[1], v[2], 2);
            }
            ++iter;
        }
    }
    void inverse(double inv[3][3]) const {
        const auto& a = ax;
        const double det = a[0][0] * a[1][1] * a[2][2] + a[0][1] * a[1][2] * a[2][0] +
                           a[0][2] * a[1][0] * a[2][1] - a[0][2] * a[1][1] * a[2][0] -
                           a[0][1] * a[1][0] * a[2][2] - a[0][0] * a[1][2] * a[2][1];
        inv[0][0] = a[1][1] * a[2][2] - a[1][2] * a[2][1];
        inv[0][1] = -(a[0][1] * a[2][2] - a[0][2] * a[2][1]);
        inv[0][2] = a[0][1] * a[1][2] - a[0][2] * a[1][1];
        inv[1][0] = -(a[1][0] * a[2][2] - a[1][2] * a[2][0]);
        inv[1][1] = a[0][0] * a[2][2] - a[0][2] * a[2][0];
        inv[1][2] = -(a[0][0] * a[1][2] - a[0][2] * a[1][0]);
        inv[2][0] = a[1][0] * a[2][1] - a[1][1] * a[2][0];
        inv[2][1] = -(a[0][0] * a[2][1] - a[0][1] * a[2][0]);
        inv[2][2] = a[0][0] * a[1][1] - a[0][1] * a[1][0];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) inv[i][j] /= det;
    }
};

}  // namespace

void cell_translations(const double* axis9, const double* pos_c, int n, double cutoff,
                       std::vector<double>& trans, std::vector<double>& pos_out, CellInfo* info) {
    Cell c;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) c.ax[i][j] = axis9[3 * i + j];
    c.metric();
    pos_out.assign(pos_c, pos_c + 3 * (size_t)n);  // 3 x n (component-major)
    if (c.rsum > 0) {
        c.refine();
        double inv[3][3];
        c.inverse(inv);
        for (int j = 0; j < n; ++j) {
            const double pc[3] = {pos_c[j], pos_c[n + j], pos_c[2 * n + j]};
            double f[3];
            for (int a = 0; a < 3; ++a)
                f[a] = inv[a][0] * pc[0] + inv[a][1] * pc[1] + inv[a][2] * pc[2];
            for (int a = 0; a < 3; ++a) f[a] -= std::floor(f[a]);
            double out[3];
            c.cart(f[0], f[1], f[2], out);
            pos_out[j] = out[0], pos_out[n + j] = out[1], pos_out[2 * n + j] = out[2];
        }
    }
    const int m0 = int(std::ceil(cutoff / c.dist(1, 0, 0)) + 1);
    const int m1 = int(std::ceil(cutoff / c.dist(0, 1, 0)) + 1);
    const int m2 = int(std::ceil(cutoff / c.dist(0, 0, 1)) + 1);
    double maxlen = 0.0;
    for (int i = -1; i < 2; ++i)
        for (int j = -1; j < 2; ++j)
            for (int k = -1; k < 2; ++k) {
                const double d = c.dist(i, j, k);
                if (maxlen < d) maxlen = d;
            }
    trans.clear();
    if (info) {
        info->mexp[0] = m0, info->mexp[1] = m1, info->mexp[2] = m2;
        info->tindex.assign((size_t)(2 * m0 + 1) * (2 * m1 + 1) * (2 * m2 + 1), -1);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) info->axis[3 * a + b] = c.ax[a][b];
    }
    int t = 0;
    for (int i = -m0; i < m0 + 1; ++i)
        for (int j = -m1; j < m1 + 1; ++j)
            for (int k = -m2; k < m2 + 1; ++k) {
                double v[3];
                c.cart(i, j, k, v);
                const double d = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                if (d < maxlen + cutoff) {
                    trans.insert(trans.end(), v, v + 3);
                    if (info)
                        info->tindex[((size_t)(i + m0) * (2 * m1 + 1) + (j + m1)) * (2 * m2 + 1) +
                                     (k + m2)] = t;
                    ++t;
                }
            }
}

}  // namespace plm
