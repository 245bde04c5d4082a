// Polynomial columns of X by factor-derivative products: host tables and the
// per-sub-batch sequence (factor values and monomials, factor derivatives,
// then per group of equal-size structures the entry products on MFMA and the
// sum of each column's entries into the X rows, kernels_polyz.hip). The
// column order and the products follow Model::model_polynomial /
// model_order{2,3} (src/cxx/src/compute/model.cpp:159-267).
#include <algorithm>
#include <array>
#include <map>
#include <stdexcept>
#include <vector>

#include "capi_internal.hpp"
#include "pair_data.hpp"
#include "profile.hpp"

namespace plm {

const PolyTables& ModelHandle::poly() {
    if (poly_tables) return *poly_tables;
    auto pt = std::make_unique<PolyTables>();
    const Model& M = host;
    const int nlin = M.n_linear(), F = M.n_features(), nt = M.n_type();
    std::vector<std::vector<int>> cols;
    for (const auto& c : M.comb2()) cols.push_back({c[0], c[1]});
    for (const auto& c : M.comb3()) cols.push_back({c[0], c[1], c[2]});
    if ((int)cols.size() != F - nlin) throw std::logic_error("polynomial column count");
    std::map<int, int> phi;
    for (const auto& c : cols)
        for (int f : c) phi.emplace(f, 0);
    int K = 0;
    for (auto& kv : phi) kv.second = K++;
    bool deg2 = false;
    for (const auto& c : cols) deg2 = deg2 || c.size() == 3;
    std::vector<int> ma, mb;
    for (int a = 0; a < K; ++a) ma.push_back(a), mb.push_back(-1);
    std::map<std::pair<int, int>, int> pair_u;
    if (deg2)
        for (int a = 0; a < K; ++a)
            for (int b = a; b < K; ++b) {
                pair_u[{a, b}] = (int)ma.size();
                ma.push_back(a), mb.push_back(b);
            }
    std::vector<int> cphi(cols.size() * 3, -1), cu(cols.size() * 3, -1);
    for (size_t c = 0; c < cols.size(); ++c) {
        std::vector<int> q;
        for (int f : cols[c]) q.push_back(phi.at(f));
        for (size_t s = 0; s < q.size(); ++s) {
            cphi[3 * c + s] = q[s];
            std::vector<int> rest;
            for (size_t k = 0; k < q.size(); ++k)
                if (k != s) rest.push_back(q[k]);
            if (rest.size() == 1) cu[3 * c + s] = rest[0];
            else cu[3 * c + s] = pair_u.at({std::min(rest[0], rest[1]), std::max(rest[0], rest[1])});
        }
    }
    std::vector<int> floc((size_t)nt * std::max(K, 1), -1);
    for (int t = 0; t < nt; ++t) {
        const auto& lg = M.type(t).linear_global;
        for (size_t k = 0; k < lg.size(); ++k) {
            auto it = phi.find(lg[k]);
            if (it != phi.end()) floc[(size_t)t * K + it->second] = (int)k;
        }
    }
    // entries: the distinct (factor, monomial) pairs of the columns' slots,
    // sorted by factor then monomial; col_pos maps (column, slot) to its entry
    std::map<std::pair<int, int>, int> ent_id;
    for (size_t c = 0; c < cols.size(); ++c)
        for (int sl = 0; sl < 3; ++sl)
            if (cphi[3 * c + sl] >= 0) ent_id.emplace(std::make_pair(cphi[3 * c + sl], cu[3 * c + sl]), 0);
    std::vector<int> ent_u, col_pos(cols.size() * 3, -1);
    pt->ent_off.assign((size_t)K + 1, 0);
    for (auto& kv : ent_id) {
        kv.second = (int)ent_u.size();
        ent_u.push_back(kv.first.second);
        ++pt->ent_off[(size_t)kv.first.first + 1];
    }
    for (int f = 0; f < K; ++f) pt->ent_off[f + 1] += pt->ent_off[f];
    for (size_t c = 0; c < cols.size(); ++c)
        for (int sl = 0; sl < 3; ++sl)
            if (cphi[3 * c + sl] >= 0) col_pos[3 * c + sl] = ent_id.at({cphi[3 * c + sl], cu[3 * c + sl]});
    std::vector<int> zt;
    for (int f = 0; f < K; ++f)
        for (int e0 = pt->ent_off[f]; e0 < pt->ent_off[f + 1]; e0 += 128)
            zt.insert(zt.end(), {f, e0, std::min(e0 + 128, pt->ent_off[f + 1])});
    PolyDev& d = pt->dev;
    d.n_ztiles = (int)zt.size() / 3;
    d.ztiles = pt->store.upload(zt.data(), zt.size() * 4);
    std::vector<int> cp4(cols.size() * 4, -1);
    for (size_t c = 0; c < cols.size(); ++c)
        for (int sl = 0; sl < 3; ++sl) cp4[4 * c + sl] = col_pos[3 * c + sl];
    d.col_pos4 = reinterpret_cast<const int4*>(pt->store.upload(cp4.data(), cp4.size() * 4));
    d.E = (int)ent_u.size();
    d.ent_u = pt->store.upload(ent_u.data(), ent_u.size() * 4);
    d.col_pos = pt->store.upload(col_pos.data(), col_pos.size() * 4);
    d.K = K;
    d.U = (int)ma.size();
    d.n_cols = (int)cols.size();
    d.c0 = nlin;
    d.fac_local = pt->store.upload(floc.data(), floc.size() * 4);
    d.col_phi = pt->store.upload(cphi.data(), cphi.size() * 4);
    d.col_u = pt->store.upload(cu.data(), cu.size() * 4);
    d.mono_a = pt->store.upload(ma.data(), ma.size() * 4);
    d.mono_b = pt->store.upload(mb.data(), mb.size() * 4);
    pt->enabled = d.n_cols > 0 && K > 0;
    poly_tables = std::move(pt);
    return *poly_tables;
}

size_t poly_bytes(ModelHandle& m, const BatchHandle& b, int s0, int s1, bool fp64) {
    const PolyTables& P = m.poly();
    if (!P.enabled) return 0;
    size_t n = 0;
    for (int s = s0; s < s1; ++s) {
        const size_t N = b.atom_off[s + 1] - b.atom_off[s];
        n += N * P.dev.K * (6 + 3 * N) + N * (P.dev.K + P.dev.U);
    }
    return n * (fp64 ? 8 : 4);
}

// LDS per workgroup of the factor-derivative kernel. Every pass walks all of
// an atom's pairs with one thread per factor of the chunk, so the factors are
// split into as few passes as the LDS holds, in chunks of equal size.
constexpr size_t kDerivLds = 150 * 1024;

std::vector<long> poly_d_offsets(ModelHandle& m, const BatchHandle& b, int s0, int s1) {
    const PolyDev& pd = m.poly().dev;
    std::vector<long> doff(s1 - s0 + 1, 0);
    for (int s = s0; s < s1; ++s) {
        const long N = b.atom_off[s + 1] - b.atom_off[s];
        doff[s - s0 + 1] = doff[s - s0] + N * pd.K * (6 + 3 * N);
    }
    return doff;
}

void build_poly_rows(ModelHandle& m, BatchHandle& b, int s0, int s1, bool fp64, void* X, int ld,
                     const double* roww, double* xe_sum, double* xe_sq, XWork& work,
                     long pair_base, const PairRec& R, int w_stride, int x_stride,
                     const long* doff_dev) {
    const PolyTables& P = m.poly();
    const PolyDev& pd = P.dev;
    Device& dev = device();
    hipStream_t st = dev.stream;
    const int elem = fp64 ? 8 : 4;
    const int a0 = b.atom_off[s0], na = b.atom_off[s1] - a0;
    const BatchDev bd = b.view();
    work.xf.alloc((size_t)na * pd.K * elem);
    work.m.alloc((size_t)na * pd.U * elem);
    // D offsets per structure of the sub-batch (uploaded here unless the caller
    // holds them on the device already, as a captured sequence must)
    const std::vector<long> doff = poly_d_offsets(m, b, s0, s1);
    int nmax = 0;
    for (int s = s0; s < s1; ++s) nmax = std::max(nmax, b.atom_off[s + 1] - b.atom_off[s]);
    const int rd_max = 6 + 3 * nmax;
    work.d.alloc((size_t)doff.back() * elem);
    if (!doff_dev) {
        work.doff.alloc(doff.size() * sizeof(long));
        // a pageable-memory copy returns once the source has been read
        check_hip(hipMemcpyAsync(work.doff.ptr, doff.data(), doff.size() * sizeof(long),
                                 hipMemcpyHostToDevice, st), "copy");
        doff_dev = (const long*)work.doff.ptr;
    }
    {
        ProfScope ps("poly_deriv", st);
        launch_poly_factor(st, fp64, m.md, pd, bd, a0, na, work.x.ptr, x_stride, work.xf.ptr,
                           work.m.ptr);
        // factor chunks whose per-atom rows fit kDerivLds bytes of LDS
        // (at most 128, so that the 256-thread workgroup has at least two
        // threads per factor)
        const int kc_max = std::max(
            1, std::min({pd.K, 128, (int)(kDerivLds / ((size_t)rd_max * elem))}));
        const int passes = (pd.K + kc_max - 1) / kc_max;
        const int kc = (pd.K + passes - 1) / passes;
        if ((size_t)kc * rd_max * elem > 160 * 1024)
            throw std::runtime_error("structure too large for the factor-derivative rows");
        for (int k0 = 0; k0 < pd.K; k0 += kc)
            launch_poly_deriv(st, fp64, m.md, pd, bd, a0, na, pair_base, work.pd.ptr, R,
                              work.w.ptr, w_stride, k0, std::min(kc, pd.K - k0), rd_max,
                              work.d.ptr, s0, doff_dev);
    }
    check_hip(hipGetLastError(), "poly launch");
    // per group of equal-size structures: the entry products Z[s][r][e] (at
    // most z_budget bytes), then the rows of the polynomial columns from them
    const size_t z_budget = 8ull << 30;
    const int E = pd.E;
    int s = s0;
    while (s < s1) {
        const int N = b.atom_off[s + 1] - b.atom_off[s];
        const long RD = 6 + 3 * N;
        const size_t y_one = (size_t)RD * E;
        int g = 1;
        while (s + g < s1 && b.atom_off[s + g + 1] - b.atom_off[s + g] == N &&
               (size_t)(g + 1) * y_one * elem <= z_budget)
            ++g;
        work.z.alloc(y_one * g * elem);
        {
            ProfScope ps("poly_z", st);
            launch_poly_z(st, fp64, pd, bd, s, g, s0, a0, (int)RD, work.d.ptr, doff_dev,
                          work.m.ptr, work.z.ptr, (long)y_one);
        }
        check_hip(hipGetLastError(), "polynomial entry launch");
        {
            ProfScope ps("poly_rows", st);
            launch_poly_rows(st, fp64, pd, bd, s, g, a0, (int)RD, work.xf.ptr, work.z.ptr,
                             (long)y_one, X, ld, b.row_off[s0], roww, xe_sum, xe_sq);
        }
        check_hip(hipGetLastError(), "polynomial column launch");
        s += g;
    }
}

}  // namespace plm
