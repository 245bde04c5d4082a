// Evaluation (PolymlpEval) orchestration: potential tables, half neighbour
// list, order parameters, per-atom prod_sum and the half-pair loop.
#include <algorithm>
#include <array>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "capi_internal.hpp"
#include "profile.hpp"

namespace plm {

void upload_potential(ModelHandle& m) {
    const Model& M = m.host;
    if (!M.has_potential()) throw std::runtime_error("coefficients not set");
    m.pot_store.bufs.clear();
    const int nt = M.n_type(), ntp = M.n_tp();
    const bool gt = M.input().gtinv;
    m.max_heads = 0, m.max_lin = 0;
    // the contraction deals each atom's terms to one wave, or to four waves
    // (sharing the atom's LDS data) when a wave would hold more than 128 terms
    // per lane; PLM_PRODSUM_LANES=64|256 forces one
    long max_terms = 0;
    for (int t = 0; t < nt; ++t) {
        long n = 0;
        for (const auto& h : M.type(t).pot) n += (long)h.size();
        max_terms = std::max(max_terms, n);
    }
    int lanes = max_terms > 128L * 64 ? 256 : 64;
    if (const char* e = std::getenv("PLM_PRODSUM_LANES")) {
        if (!std::strcmp(e, "64")) lanes = 64;
        else if (!std::strcmp(e, "256")) lanes = 256;
        else throw std::invalid_argument("PLM_PRODSUM_LANES must be 64 or 256");
    }
    m.ed.pq_lanes = lanes;
    // The heads of each type in the order of the half-pair loop: grouped by type
    // pair, within one ordered by Y key (the per-lane pair loop generates Y_lm
    // key by key). The contraction writes head k of an atom at its position
    // rank[t][k] in that order, so the pair loop reads both atoms' sums in
    // address order. Entries: (radial n_id, ylm key, head k, head seen from
    // atom j, sign, m == 0).
    using HeadEntry = std::array<int, 6>;
    std::vector<std::vector<std::vector<HeadEntry>>> by_tp_t(nt);
    std::vector<std::vector<int>> order(nt);
    m.head_rank.assign(nt, {});
    for (int t = 0; t < nt; ++t) {
        const auto& T = M.type(t);
        auto& by_tp = by_tp_t[t];
        by_tp.assign(ntp, {});
        if (gt) {
            for (size_t k = 0; k < T.noconj.size(); ++k) {
                const auto& a = T.nlm[T.noconj[k]];
                by_tp[a.tp].push_back({a.n_id, a.ylm_key, (int)k, a.jlocal_noconj,
                                       a.sign_j < 0 ? -1 : 1, a.m == 0 ? 1 : 0});
            }
        } else {
            for (size_t k = 0; k < T.ntp.size(); ++k)
                by_tp[T.ntp[k][2]].push_back({T.ntp[k][1], -1, (int)k, T.ntp[k][3], 1, 0});
        }
        const int n_heads = (int)T.pot.size();
        auto& rank = m.head_rank[t];
        rank.assign(n_heads, -1);
        for (int tp = 0; tp < ntp; ++tp) {
            if (gt)
                std::stable_sort(by_tp[tp].begin(), by_tp[tp].end(),
                                 [](const HeadEntry& a, const HeadEntry& b) { return a[1] < b[1]; });
            for (const auto& e : by_tp[tp]) {
                if (e[2] < 0 || e[2] >= n_heads || rank[e[2]] >= 0)
                    throw std::invalid_argument("evaluation heads do not match the potential");
                rank[e[2]] = (int)order[t].size();
                order[t].push_back(e[2]);
            }
        }
        if ((int)order[t].size() != n_heads)
            throw std::invalid_argument("evaluation heads do not match the potential");
    }
    // the type of atom j for an atom of type t and a type pair
    auto partner = [&](int t, int tp) {
        int found = -1;
        for (int t2 = 0; t2 < nt; ++t2)
            if (M.type_pair(t, t2) == tp) {
                if (found >= 0) throw std::invalid_argument("type pair with two partner types");
                found = t2;
            }
        return found;
    };
    for (int t = 0; t < nt; ++t) {
        const auto& T = M.type(t);
        EvalTypeDev& E = m.ed.t[t];
        E.n_heads = (int)T.pot.size();
        m.max_heads = std::max(m.max_heads, E.n_heads);
        m.max_lin = std::max(m.max_lin, (int)T.features.size());
        // the terms by head position
        std::vector<int> off(E.n_heads + 1, 0), pd, pf;
        std::vector<double> ce, cf;
        for (int r = 0; r < E.n_heads; ++r) {
            for (const auto& pt : T.pot[order[t][r]]) {
                const auto& k3 = T.prod_deriv[pt.prod];
                const auto& k2 = T.prod_features[pt.prod_features];
                if (k3.size() > 3 || k2.size() > 2) throw std::invalid_argument("term order too high");
                for (int a = 0; a < 3; ++a) pd.push_back(a < (int)k3.size() ? k3[a] : -1);
                for (int a = 0; a < 2; ++a) pf.push_back(a < (int)k2.size() ? k2[a] : -1);
                ce.push_back(pt.coeff_e), cf.push_back(pt.coeff_f);
            }
            off[r + 1] = (int)ce.size();
        }
        // per type pair its heads and the first head of every Y key; hi and hj
        // are head positions of atoms i and j
        const int nkey = m.md.n_lmh;
        std::vector<int> hoff(ntp + 1, 0), koff((size_t)ntp * (nkey + 1), 0), nid, ylm, hi, hj;
        std::vector<float> sj;
        std::vector<unsigned char> m0;
        for (int tp = 0; tp < ntp; ++tp) {
            const auto& heads = by_tp_t[t][tp];
            const int t2 = heads.empty() ? -1 : partner(t, tp);
            if (!heads.empty() && t2 < 0) throw std::invalid_argument("type pair without partner type");
            int* ko = koff.data() + (size_t)tp * (nkey + 1);
            for (const auto& e : heads) {
                if (gt && (e[1] < 0 || e[1] >= nkey)) throw std::invalid_argument("head Y key range");
                if (e[3] < 0 || e[3] >= (int)m.head_rank[t2].size())
                    throw std::invalid_argument("head of atom j out of range");
                if (gt) ++ko[e[1] + 1];
                nid.push_back(e[0]), ylm.push_back(e[1]);
                hi.push_back(m.head_rank[t][e[2]]), hj.push_back(m.head_rank[t2][e[3]]);
                sj.push_back((float)e[4]), m0.push_back((unsigned char)e[5]);
            }
            ko[0] = hoff[tp];
            for (int k = 0; k < nkey; ++k) ko[k + 1] += ko[k];
            hoff[tp + 1] = (int)nid.size();
        }
        auto upv = [&](const auto& v) {
            using V = typename std::decay_t<decltype(v)>::value_type;
            return m.pot_store.upload(v.data(), v.size() * sizeof(V));
        };
        E.pot_off = upv(off);
        // interleaved lane chunks of the contraction (keys and heads fit 16 bits:
        // the order parameters and features they index are held in LDS)
        const long n_terms = (long)ce.size();
        const long iters = (n_terms + lanes - 1) / lanes;
        std::vector<int4> pq_key((size_t)std::max(1L, iters) * lanes, int4{-1, -1, -1, 0});
        std::vector<double2> pq_c(pq_key.size(), double2{0.0, 0.0});
        auto k16 = [](int v) -> unsigned {
            if (v >= 0xffff) throw std::invalid_argument("evaluation table index exceeds 16 bits");
            return v < 0 ? 0xffffu : (unsigned)v;
        };
        for (int h = 0; h < E.n_heads; ++h)
            for (int e = off[h]; e < off[h + 1]; ++e) {
                const size_t r = (size_t)(e % iters) * lanes + e / iters;
                pq_key[r] = int4{(int)(k16(pd[3 * e]) | k16(pd[3 * e + 1]) << 16),
                                 (int)(k16(pd[3 * e + 2]) | k16(h) << 16),
                                 (int)(k16(pf[2 * e]) | k16(pf[2 * e + 1]) << 16), 0};
                pq_c[r] = double2{ce[e], cf[e]};
            }
        E.pq_iters = (int)iters;
        E.pq_key = static_cast<const int4*>(upv(pq_key));
        E.pq_c = static_cast<const double2*>(upv(pq_c));
        std::vector<int4> rec(std::max<size_t>(1, nid.size()), int4{0, 0, 0, 0});
        for (size_t h = 0; h < nid.size(); ++h) {
            if (hj[h] < 0 || hj[h] >= (1 << 28)) throw std::invalid_argument("head index range");
            rec[h] = int4{nid[h], ylm[h], hi[h],
                          hj[h] | (sj[h] < 0.f ? 1 << 28 : 0) | (m0[h] ? 1 << 29 : 0)};
        }
        E.hd_off = upv(hoff);
        E.hd_koff = upv(koff);
        E.hd_rec = static_cast<const int4*>(upv(rec));
    }
    m.pot_uploaded = true;
    ++m.pot_version;
}

void BatchHandle::ensure_half() {
    if (has_half) return;
    std::vector<int> off;
    n_half = build_list(true, d_hb_off, d_hb_j, d_hb_t, d_hb_d, off);
    has_half = true;
}

// Work buffers of one evaluation; out = [energies | forces (3 per atom) | stresses (6)].
static void alloc_eval(const ModelHandle& m, const BatchHandle& b, bool fp64, DevBuf& a_buf,
                       DevBuf& pse, DevBuf& psf, DevBuf& out, size_t& n_out) {
    const int elem = fp64 ? 8 : 4;
    const size_t A = b.n_atoms;
    a_buf.alloc(A * m.max_nlm * 2 * elem);
    pse.alloc(A * std::max(1, m.max_heads) * 2 * elem);
    psf.alloc(A * std::max(1, m.max_heads) * 2 * elem);
    n_out = (size_t)b.n_st + 3 * A + 6 * (size_t)b.n_st;
    out.alloc(n_out * 8);
}

// The kernels of one evaluation on st: out is cleared by a kernel (no memset
// node when captured), then filled by the half-pair loop. bd carries the full
// neighbour list (order parameters); hb_* the half list (pair loop).
void enqueue_eval_dev(ModelHandle& m, const BatchDev& bd, const int* hb_off, const int* hb_j,
                      const double* hb_d, int n_half, bool fp64, void* a_buf, void* pse,
                      void* psf, double* out, size_t n_out, hipStream_t st) {
    const int A = bd.n_atoms_total;
    launch_fill_zero(st, out, n_out);
    double* d_e = out;
    double* d_f = d_e + bd.n_st;
    double* d_s = d_f + 3 * (size_t)A;
    const int pair_model = m.host.input().gtinv ? 0 : 1;
    {
        ProfScope ps("eval_anlm", st);
        launch_anlm(st, fp64, m.md, bd, a_buf, m.max_nlm, 0, A, m.max_noconj, 1);
    }
    check_hip(hipGetLastError(), "eval anlm launch");
    {
        ProfScope ps("eval_prodsum", st);
        launch_prodsum(st, fp64, m.md, m.ed, bd, a_buf, m.max_nlm, pse, psf,
                       std::max(1, m.max_heads), m.max_lin, pair_model);
    }
    check_hip(hipGetLastError(), "prodsum launch");
    {
        ProfScope ps("eval_pair", st);
        launch_pair(st, fp64, m.md, m.ed, bd, bd.atom_st, hb_off, hb_j, hb_d, n_half, pse, psf,
                    std::max(1, m.max_heads), d_e, d_f, d_s, pair_model);
    }
    check_hip(hipGetLastError(), "pair launch");
}

static void enqueue_eval(ModelHandle& m, BatchHandle& b, bool fp64, void* a_buf, void* pse,
                         void* psf, double* out, size_t n_out, hipStream_t st) {
    enqueue_eval_dev(m, b.view(), (const int*)b.d_hb_off.ptr, (const int*)b.d_hb_j.ptr,
                     (const double*)b.d_hb_d.ptr, (int)b.n_half, fp64, a_buf, pse, psf, out,
                     n_out, st);
}

static void copy_out(const BatchHandle& b, const DevBuf& out, size_t n_out, hipStream_t st,
                     double* energies, double* forces, double* stresses) {
    const size_t A = b.n_atoms;
    std::vector<double> h(n_out);
    check_hip(hipMemcpyAsync(h.data(), out.ptr, n_out * 8, hipMemcpyDeviceToHost, st), "copy");
    check_hip(hipStreamSynchronize(st), "sync");
    if (energies) std::memcpy(energies, h.data(), (size_t)b.n_st * 8);
    if (forces) std::memcpy(forces, h.data() + b.n_st, 3 * A * 8);
    if (stresses) std::memcpy(stresses, h.data() + b.n_st + 3 * A, 6 * (size_t)b.n_st * 8);
}

static void prepare_eval(ModelHandle& m, BatchHandle& b) {
    if (!m.built) m.build_device();
    if (!m.pot_uploaded) upload_potential(m);
    b.ensure_half();
}

void eval_batch(ModelHandle& m, BatchHandle& b, bool fp64, double* energies, double* forces,
                double* stresses) {
    prepare_eval(m, b);
    hipStream_t st = device().stream;
    std::lock_guard<std::mutex> lock(m.eval_mutex);
    size_t n_out = 0;
    alloc_eval(m, b, fp64, m.ev_a, m.ev_pse, m.ev_psf, m.ev_out, n_out);
    enqueue_eval(m, b, fp64, m.ev_a.ptr, m.ev_pse.ptr, m.ev_psf.ptr, (double*)m.ev_out.ptr, n_out,
                 st);
    copy_out(b, m.ev_out, n_out, st, energies, forces, stresses);
}

EvalGraph::~EvalGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
}

void eval_graph_capture(EvalGraph& g) {
    ModelHandle& m = *g.m;
    BatchHandle& b = *g.b;
    prepare_eval(m, b);
    hipStream_t st = device().stream;
    alloc_eval(m, b, g.fp64, g.a_buf, g.pse, g.psf, g.out, g.n_out);
    // one eager pass first: kernel attributes are set and every buffer exists
    enqueue_eval(m, b, g.fp64, g.a_buf.ptr, g.pse.ptr, g.psf.ptr, (double*)g.out.ptr, g.n_out, st);
    check_hip(hipStreamSynchronize(st), "sync");
    {
        ProfPause pause;
        check_hip(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "begin capture");
        try {
            enqueue_eval(m, b, g.fp64, g.a_buf.ptr, g.pse.ptr, g.psf.ptr, (double*)g.out.ptr,
                         g.n_out, st);
        } catch (...) {
            hipGraph_t tmp = nullptr;
            (void)hipStreamEndCapture(st, &tmp);
            if (tmp) (void)hipGraphDestroy(tmp);
            throw;
        }
        check_hip(hipStreamEndCapture(st, &g.graph), "end capture");
    }
    check_hip(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0), "graph instantiate");
    g.pot_version = m.pot_version;
}

void eval_graph_run(EvalGraph& g, int steps, double* energies, double* forces, double* stresses) {
    if (steps < 1) throw std::invalid_argument("steps must be positive");
    if (!g.m->pot_uploaded || g.m->pot_version != g.pot_version)
        throw std::invalid_argument("model coefficients changed since the graph was captured");
    hipStream_t st = device().stream;
    for (int k = 0; k < steps; ++k) check_hip(hipGraphLaunch(g.exec, st), "graph launch");
    copy_out(*g.b, g.out, g.n_out, st, energies, forces, stresses);
}

void eval_single(ModelHandle& m, const double* axis, const double* pos_c, const int* types, int n,
                 bool fp64, double* energy, double* forces, double* stress) {
    if (!m.built) m.build_device();
    BatchHandle b(&m, m.host.input().cutoff, 1, axis, pos_c, types, &n);
    eval_batch(m, b, fp64, energy, forces, stress);
}

}  // namespace plm
