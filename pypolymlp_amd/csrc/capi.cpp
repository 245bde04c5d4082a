// C-ABI implementation: model and batch device tables, X rows, fused X^T X.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/polymlp_amd.h"
#include "capi_internal.hpp"
#include "pair_data.hpp"
#include "cell.hpp"
#include "kernels.hpp"
#include "model.hpp"
#include "profile.hpp"

namespace plm {

thread_local std::string g_last_error;

void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}

DevBuf::~DevBuf() { release(); }
void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr, bytes = 0;
}
void DevBuf::alloc(size_t n) {
    if (n <= bytes && ptr) return;
    release();
    if (n == 0) n = 16;
    check_hip(hipMalloc(&ptr, n), "hipMalloc");
    bytes = n;
}

Device& device() {
    static Device d;
    if (!d.init) {
        int n = 0;
        check_hip(hipGetDeviceCount(&n), "hipGetDeviceCount");
        if (n < 1) throw DeviceError("no HIP device");
        check_hip(hipSetDevice(0), "hipSetDevice");
        check_hip(hipStreamCreate(&d.stream), "hipStreamCreate");
        d.init = true;
    }
    return d;
}

// ---------------------------------------------------------------------------
// model tables
// ---------------------------------------------------------------------------
ModelHandle::ModelHandle(const ModelInput& in) : host(in) {}

template <class V>
const typename V::value_type* up(ModelHandle& h, const V& v) {
    return h.store.upload(v.data(), v.size() * sizeof(typename V::value_type));
}

void ModelHandle::build_device() {
    const Model& M = host;
    const int nt = M.n_type(), ntp = M.n_tp(), nfn = M.n_fn();
    const bool gt = M.input().gtinv;
    md = ModelDev{};
    md.n_type = nt, md.n_fn = nfn, md.n_tp = ntp;
    md.max_l = gt ? M.input().max_l : -1;
    md.n_lmh = gt ? M.n_lm_half() : 1;
    md.cutoff = M.input().cutoff;
    md.n_features = M.n_features();
    md.n_linear = M.n_linear();
    if (nt > kMaxTypes) throw std::invalid_argument("too many atom types");
    if (md.n_lmh > kMaxLmHalf) throw std::invalid_argument("max_l too large");
    for (int a = 0; a < nt; ++a)
        for (int b = 0; b < nt; ++b) md.tp_of[a * nt + b] = M.type_pair(a, b);
    std::vector<double> rad((size_t)ntp * nfn * 2, 0.0);
    std::vector<int> rcount(ntp);
    for (int tp = 0; tp < ntp; ++tp) {
        const auto& ns = M.tp_to_n()[tp];
        rcount[tp] = (int)ns.size();
        for (size_t k = 0; k < ns.size(); ++k) {
            rad[((size_t)tp * nfn + k) * 2] = M.input().params[ns[k]][0];
            rad[((size_t)tp * nfn + k) * 2 + 1] = M.input().params[ns[k]][1];
        }
    }
    md.rad = up(*this, rad);
    md.rad_count = up(*this, rcount);
    max_nlm = 0, max_noconj = 0;
    for (int t = 0; t < nt; ++t) {
        const auto& T = M.type(t);
        TypeDev& D = md.t[t];
        D = TypeDev{};
        // order-parameter components
        std::vector<int> src, tpo(ntp + 1, 0), head, nid, ylmk;
        std::vector<float> sign;
        std::vector<unsigned char> conj;
        std::vector<std::vector<std::array<int, 3>>> by_tp(ntp);
        if (gt) {
            D.n_nlm = (int)T.nlm.size();
            D.n_noconj = (int)T.noconj.size();
            for (const auto& a : T.nlm) {
                if (a.conj) {
                    src.push_back(T.local_to_noconj[a.ilocal_conj]);
                    sign.push_back((float)a.cc);
                    conj.push_back(1);
                } else {
                    src.push_back(T.local_to_noconj[a.ilocal]);
                    sign.push_back(1.f);
                    conj.push_back(0);
                }
            }
            for (size_t k = 0; k < T.noconj.size(); ++k) {
                const auto& a = T.nlm[T.noconj[k]];
                by_tp[a.tp].push_back({(int)k, a.n_id, a.ylm_key});
            }
        } else {
            D.n_nlm = D.n_noconj = (int)T.ntp.size();
            for (int k = 0; k < D.n_nlm; ++k) {
                src.push_back(k), sign.push_back(1.f), conj.push_back(0);
                by_tp[T.ntp[k][2]].push_back({k, T.ntp[k][1], -1});
            }
        }
        for (int tp = 0; tp < ntp; ++tp) {
            tpo[tp + 1] = tpo[tp] + (int)by_tp[tp].size();
            for (const auto& e : by_tp[tp]) head.push_back(e[0]), nid.push_back(e[1]), ylmk.push_back(e[2]);
        }
        max_nlm = std::max(max_nlm, D.n_nlm);
        max_noconj = std::max(max_noconj, D.n_noconj);
        D.nlm_src = up(*this, src);
        D.nlm_sign = up(*this, sign);
        D.nlm_conj = up(*this, conj);
        D.nc_tp_off = up(*this, tpo);
        D.nc_tp_head = up(*this, head);
        D.nc_tp_nid = up(*this, nid);
        D.nc_tp_ylm = up(*this, ylmk);
        // linear feature values
        const int L = (int)T.features.size();
        D.n_lin = L;
        std::vector<int> fvo(L + 1, 0), fvk;
        std::vector<double> fvc;
        for (int f = 0; f < L; ++f) {
            for (const auto& pc : T.feat_val[f]) {
                const auto& keys = T.prod[pc.first];
                if (keys.size() > 4) throw std::invalid_argument("invariant order above 4");
                for (int k = 0; k < 4; ++k) fvk.push_back(k < (int)keys.size() ? keys[k] : -1);
                fvc.push_back(pc.second);
            }
            fvo[f + 1] = (int)fvc.size();
        }
        // products evaluated one by one and dropped when |Re| <= 1e-20 when the
        // type has more than 1000 of them (Features::compute_features,
        // polymlp_features.cpp:184-193); the pair form has no such rule
        D.fv_skip = (gt && T.prod.size() > 1000) ? 1 : 0;
        D.fv_off = up(*this, fvo);
        D.fv_coeff = up(*this, fvc);
        D.fv_keys = up(*this, fvk);
        // derivative weights W'[f, noconj head]
        std::vector<int> wzo((size_t)L * ntp + 1, 0), wzn, wzy, wto(1, 0), wtk;
        std::vector<double> wtc;
        std::vector<unsigned char> wtj;
        for (int f = 0; f < L; ++f) {
            // (tp, head) -> terms
            std::map<std::pair<int, int>, std::vector<std::tuple<double, int, unsigned char>>> nz;
            for (const auto& d : T.feat_deriv[f]) {
                int h0, tp;
                double c = d.coeff;
                unsigned char cj = 0;
                if (gt) {
                    const auto& a = T.nlm[d.head];
                    if (a.conj) {
                        h0 = T.local_to_noconj[a.ilocal_conj];
                        c *= a.cc;
                        cj = 1;
                    } else {
                        h0 = T.local_to_noconj[d.head];
                    }
                    tp = a.tp;
                } else {
                    h0 = d.head;
                    tp = T.ntp[d.head][2];
                }
                nz[{tp, h0}].push_back({c, d.prod, cj});
            }
            int cur_tp = 0;
            for (const auto& kv : nz) {
                const int tp = kv.first.first, h0 = kv.first.second;
                while (cur_tp < tp) wzo[(size_t)f * ntp + (++cur_tp)] = (int)wzn.size();
                if (gt) {
                    const auto& a = T.nlm[T.noconj[h0]];
                    wzn.push_back(a.n_id), wzy.push_back(a.ylm_key);
                } else {
                    wzn.push_back(T.ntp[h0][1]), wzy.push_back(-1);
                }
                for (const auto& term : kv.second) {
                    const auto& keys = T.prod_deriv[std::get<1>(term)];
                    if (keys.size() > 3) throw std::invalid_argument("invariant order above 4");
                    for (int k = 0; k < 3; ++k) wtk.push_back(k < (int)keys.size() ? keys[k] : -1);
                    wtc.push_back(std::get<0>(term));
                    wtj.push_back(std::get<2>(term));
                }
                wto.push_back((int)wtc.size());
            }
            while (cur_tp < ntp) wzo[(size_t)f * ntp + (++cur_tp)] = (int)wzn.size();
        }
        D.wz_off = up(*this, wzo);
        D.wz_nid = up(*this, wzn);
        D.wz_ylm = up(*this, wzy);
        D.wz_term_off = up(*this, wto);
        D.wt_coeff = up(*this, wtc);
        D.wt_keys = up(*this, wtk);
        D.wt_conj = up(*this, wtj);
        {
            // interleaved W' terms: items sorted by decreasing term count
            const int nnz = (int)wzn.size();
            std::vector<int> cnt(nnz), order(nnz);
            for (int g = 0; g < nnz; ++g) cnt[g] = wto[g + 1] - wto[g];
            std::iota(order.begin(), order.end(), 0);
            std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cnt[x] > cnt[y]; });
            const int nblk = (nnz + 63) / 64;
            std::vector<int> boff(nblk + 1, 0), item((size_t)nblk * 64, -1), keys;
            std::vector<double> cc;
            for (int bk = 0; bk < nblk; ++bk) {
                int rows = 0;
                for (int l = 0; l < 64 && bk * 64 + l < nnz; ++l) rows = std::max(rows, cnt[order[bk * 64 + l]]);
                boff[bk + 1] = boff[bk] + rows * 64;
                keys.resize((size_t)boff[bk + 1] * 4, -1);
                cc.resize((size_t)boff[bk + 1] * 2, 0.0);
                for (int l = 0; l < 64 && bk * 64 + l < nnz; ++l) {
                    const int g = order[bk * 64 + l];
                    item[(size_t)bk * 64 + l] = g;
                    for (int j = 0; j < cnt[g]; ++j) {
                        const int e = wto[g] + j;
                        const size_t r = (size_t)boff[bk] + (size_t)j * 64 + l;
                        int nf = 0;
                        for (int k = 0; k < 3; ++k) {
                            keys[r * 4 + k] = wtk[(size_t)e * 3 + k];
                            nf += wtk[(size_t)e * 3 + k] >= 0;
                        }
                        w_maxf = std::max(w_maxf, nf);
                        cc[r * 2] = wtc[e];
                        cc[r * 2 + 1] = wtj[e] ? -wtc[e] : wtc[e];
                    }
                }
            }
            std::vector<float> ccf(cc.begin(), cc.end());
            D.wq_nblk = nblk;
            D.wq_boff = up(*this, boff);
            D.wq_item = up(*this, item);
            D.wq_keys = up(*this, keys);
            D.wq_cc = up(*this, cc);
            D.wq_ccf = up(*this, ccf);
        }
        {
            // interleaved value terms
            std::vector<int> cnt(L), order(L);
            for (int f = 0; f < L; ++f) cnt[f] = fvo[f + 1] - fvo[f];
            std::iota(order.begin(), order.end(), 0);
            std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cnt[x] > cnt[y]; });
            const int nblk = (L + 63) / 64;
            std::vector<int> boff(nblk + 1, 0), item((size_t)nblk * 64, -1), keys;
            std::vector<double> c;
            for (int bk = 0; bk < nblk; ++bk) {
                int rows = 0;
                for (int l = 0; l < 64 && bk * 64 + l < L; ++l) rows = std::max(rows, cnt[order[bk * 64 + l]]);
                boff[bk + 1] = boff[bk] + rows * 64;
                keys.resize((size_t)boff[bk + 1] * 4, -1);
                c.resize((size_t)boff[bk + 1], 0.0);
                for (int l = 0; l < 64 && bk * 64 + l < L; ++l) {
                    const int f = order[bk * 64 + l];
                    item[(size_t)bk * 64 + l] = f;
                    for (int j = 0; j < cnt[f]; ++j) {
                        const int e = fvo[f] + j;
                        const size_t r = (size_t)boff[bk] + (size_t)j * 64 + l;
                        int nv = 0;
                        for (int k = 0; k < 4; ++k) {
                            keys[r * 4 + k] = fvk[(size_t)e * 4 + k];
                            nv += fvk[(size_t)e * 4 + k] >= 0;
                        }
                        v_maxf = std::max(v_maxf, nv);
                        c[r] = fvc[e];
                    }
                }
            }
            std::vector<float> cf(c.begin(), c.end());
            D.fq_nblk = nblk;
            D.fq_boff = up(*this, boff);
            D.fq_item = up(*this, item);
            D.fq_keys = up(*this, keys);
            D.fq_c = up(*this, c);
            D.fq_cf = up(*this, cf);
        }
        D.lin_global = up(*this, T.linear_global);
        host_wz_off.push_back(wzo);
        host_wz_ylm.push_back(wzy);
        host_wz_nid.push_back(wzn);
        wz_max_type = std::max(wz_max_type, (int)wzn.size());
        lin_max_type = std::max(lin_max_type, L);
    }
    built = true;
}

// Column recipes and per-(type, block) tables for a block width CB.
const ColumnTables& ModelHandle::columns(int CB, int cover) {
    const int key = CB * 2 + (cover < host.n_features() ? 1 : 0);
    auto it = col_cache.find(key);
    if (it != col_cache.end()) return *it->second;
    const Model& M = host;
    const int nt = M.n_type(), F = M.n_features(), nlin = M.n_linear(), ntp = M.n_tp();
    auto ct = std::make_unique<ColumnTables>();
    std::vector<unsigned char> nfac(F);
    std::vector<std::array<int, 3>> base(F, {-1, -1, -1});
    for (int c = 0; c < F; ++c) {
        if (c < nlin) nfac[c] = 1, base[c] = {c, -1, -1};
        else if (c < nlin + (int)M.comb2().size()) {
            const auto& cb = M.comb2()[c - nlin];
            nfac[c] = 2, base[c] = {cb[0], cb[1], -1};
        } else {
            const auto& cb = M.comb3()[c - nlin - M.comb2().size()];
            nfac[c] = 3, base[c] = {cb[0], cb[1], cb[2]};
        }
    }
    // global linear id -> local feature id per type
    std::vector<std::vector<int>> g2l(nt, std::vector<int>(nlin, -1));
    for (int t = 0; t < nt; ++t) {
        const auto& lg = M.type(t).linear_global;
        for (size_t k = 0; k < lg.size(); ++k) g2l[t][lg[k]] = (int)k;
    }
    // blocks over columns [0, cover): all columns, or only the linear ones when
    // the polynomial columns are built by the factor-derivative GEMM
    const int nblk = (cover + CB - 1) / CB;
    std::vector<int> c0(nblk + 1);
    for (int b = 0; b <= nblk; ++b) c0[b] = std::min(cover, b * CB);
    std::vector<int> need_off((size_t)nt * (nblk + 1), 0), need, col_need((size_t)nt * F * 3, -1);
    std::vector<int> bw_off((size_t)nt * (nblk + 1), 0), bw_nz, need_wz;
    int need_max = 1, wz_max = 1;
    for (int t = 0; t < nt; ++t) {
        const auto& wzo = host_wz_off[t];
        for (int b = 0; b < nblk; ++b) {
            need_off[(size_t)t * (nblk + 1) + b] = (int)need.size();
            bw_off[(size_t)t * (nblk + 1) + b] = (int)bw_nz.size();
            std::set<int> s;
            for (int c = c0[b]; c < c0[b + 1]; ++c) {
                bool ok = true;
                for (int k = 0; k < nfac[c]; ++k) ok = ok && g2l[t][base[c][k]] >= 0;
                if (!ok) continue;
                for (int k = 0; k < nfac[c]; ++k) s.insert(g2l[t][base[c][k]]);
            }
            std::vector<int> lst(s.begin(), s.end());
            std::map<int, int> pos;
            int nzb = 0;
            for (size_t k = 0; k < lst.size(); ++k) {
                pos[lst[k]] = (int)k;
                const int f = lst[k];
                need.push_back(f);
                need_wz.push_back(nzb);
                for (int g = wzo[(size_t)f * ntp]; g < wzo[(size_t)(f + 1) * ntp]; ++g)
                    bw_nz.push_back(g), ++nzb;
            }
            need_max = std::max(need_max, (int)lst.size());
            wz_max = std::max(wz_max, nzb);
            for (int c = c0[b]; c < c0[b + 1]; ++c) {
                bool ok = true;
                for (int k = 0; k < nfac[c]; ++k) ok = ok && g2l[t][base[c][k]] >= 0;
                if (!ok) continue;
                for (int k = 0; k < nfac[c]; ++k)
                    col_need[((size_t)t * F + c) * 3 + k] = pos[g2l[t][base[c][k]]];
            }
        }
        need_off[(size_t)t * (nblk + 1) + nblk] = (int)need.size();
        bw_off[(size_t)t * (nblk + 1) + nblk] = (int)bw_nz.size();
    }
    // block kind (all columns linear features) and Y keys its nonzeros use
    std::vector<unsigned char> kind(nblk);
    std::vector<int> nkey(nblk, 0);
    for (int b = 0; b < nblk; ++b) {
        bool lin = true;
        for (int c = c0[b]; c < c0[b + 1]; ++c) lin = lin && nfac[c] == 1;
        kind[b] = lin ? 1 : 0;
        for (int t = 0; t < nt; ++t)
            for (int z = bw_off[(size_t)t * (nblk + 1) + b]; z < bw_off[(size_t)t * (nblk + 1) + b + 1]; ++z)
                nkey[b] = std::max(nkey[b], host_wz_ylm[t][bw_nz[z]] + 1);
        ct->nkey_max = std::max(ct->nkey_max, nkey[b]);
    }
    ColumnsDev& cd = ct->cd;
    cd.n_cols = F;
    cd.blk_kind = ct->store.upload(kind.data(), kind.size());
    cd.blk_nkey = ct->store.upload(nkey.data(), nkey.size() * 4);
    cd.nfac = ct->store.upload(nfac.data(), nfac.size());
    std::vector<int> bl((size_t)nt * F * 3, -1);
    for (int t = 0; t < nt; ++t)
        for (int c = 0; c < F; ++c)
            for (int k = 0; k < nfac[c]; ++k)
                bl[((size_t)t * F + c) * 3 + k] = g2l[t][base[c][k]];
    cd.base_local = ct->store.upload(bl.data(), bl.size() * 4);
    cd.n_blocks = nblk;
    cd.blk_c0 = ct->store.upload(c0.data(), c0.size() * 4);
    cd.need_off = ct->store.upload(need_off.data(), need_off.size() * 4);
    cd.need_feat = ct->store.upload(need.data(), need.size() * 4);
    cd.col_need = ct->store.upload(col_need.data(), col_need.size() * 4);
    // packed (key + 1) << 8 | n_id per nonzero, and per needed feature the
    // block-local nonzero offset of each type pair
    std::vector<int> bw_nk(bw_nz.size()), bw_zr(need.size() * (size_t)(ntp + 1));
    for (int t = 0; t < nt; ++t) {
        const auto& wzo = host_wz_off[t];
        for (int b = 0; b < nblk; ++b) {
            for (int z = bw_off[(size_t)t * (nblk + 1) + b]; z < bw_off[(size_t)t * (nblk + 1) + b + 1]; ++z) {
                const int g = bw_nz[z];
                if (host_wz_nid[t][g] > 255) throw std::invalid_argument("more than 256 radial functions");
                bw_nk[z] = ((host_wz_ylm[t][g] + 1) << 8) | host_wz_nid[t][g];
            }
            for (int k = need_off[(size_t)t * (nblk + 1) + b]; k < need_off[(size_t)t * (nblk + 1) + b + 1]; ++k) {
                const int f = need[k];
                for (int tp = 0; tp <= ntp; ++tp)
                    bw_zr[(size_t)k * (ntp + 1) + tp] = need_wz[k] + wzo[(size_t)f * ntp + tp] - wzo[(size_t)f * ntp];
            }
        }
    }
    ct->bw.off = ct->store.upload(bw_off.data(), bw_off.size() * 4);
    ct->bw.nz = ct->store.upload(bw_nz.data(), bw_nz.size() * 4);
    ct->bw.nk = ct->store.upload(bw_nk.data(), bw_nk.size() * 4);
    ct->bw.zr = ct->store.upload(bw_zr.data(), bw_zr.size() * 4);
    ct->need_max = need_max;
    ct->wz_max = wz_max;
    ct->CB = CB;
    auto& ref = *ct;
    col_cache[key] = std::move(ct);
    return ref;
}

// ---------------------------------------------------------------------------
// batch
// ---------------------------------------------------------------------------
BatchHandle::BatchHandle(ModelHandle* m, double cutoff_, int n_st_, const double* axes,
                         const double* pos, const int* types_in, const int* n_atoms_in)
    : model(m), cutoff(cutoff_), n_st(n_st_) {
    static std::atomic<unsigned long long> next_serial{0};
    serial = ++next_serial;
    Device& dev = device();
    atom_off.assign(n_st + 1, 0);
    for (int s = 0; s < n_st; ++s) {
        if (n_atoms_in[s] <= 0) throw std::invalid_argument("structure with no atoms");
        atom_off[s + 1] = atom_off[s] + n_atoms_in[s];
    }
    n_atoms = atom_off[n_st];
    row_off.assign(n_st + 1, 0);
    for (int s = 0; s < n_st; ++s) row_off[s + 1] = row_off[s] + 7 + 3 * n_atoms_in[s];
    std::vector<double> pos_dev((size_t)n_atoms * 3), trans;
    std::vector<int> tr_off(n_st + 1, 0), atom_st(n_atoms);
    size_t poff = 0;
    max_atoms = 0;
    for (int s = 0; s < n_st; ++s) {
        const int n = n_atoms_in[s];
        max_atoms = std::max(max_atoms, n);
        std::vector<double> tr, prev;
        cell_translations(axes + 9 * (size_t)s, pos + poff, n, cutoff, tr, prev);
        for (int j = 0; j < n; ++j)
            for (int a = 0; a < 3; ++a) pos_dev[3 * (size_t)(atom_off[s] + j) + a] = prev[(size_t)a * n + j];
        trans.insert(trans.end(), tr.begin(), tr.end());
        tr_off[s + 1] = tr_off[s] + (int)(tr.size() / 3);
        for (int j = 0; j < n; ++j) atom_st[atom_off[s] + j] = s;
        poff += 3 * (size_t)n;
    }
    types.assign(types_in, types_in + n_atoms);
    if (m)
        for (int t : types)
            if (t < 0 || t >= m->host.n_type()) throw std::invalid_argument("atom type out of range");
    d_types.alloc(n_atoms * 4);
    d_pos.alloc(pos_dev.size() * 8);
    d_trans.alloc(std::max<size_t>(trans.size(), 3) * 8);
    d_tr_off.alloc((n_st + 1) * 4);
    d_atom_st.alloc(n_atoms * 4);
    d_atom_off.alloc((n_st + 1) * 4);
    d_row_off.alloc((n_st + 1) * 4);
    hipStream_t st = dev.stream;
    check_hip(hipMemcpyAsync(d_types.ptr, types.data(), n_atoms * 4, hipMemcpyHostToDevice, st), "copy");
    check_hip(hipMemcpyAsync(d_pos.ptr, pos_dev.data(), pos_dev.size() * 8, hipMemcpyHostToDevice, st), "copy");
    if (!trans.empty())
        check_hip(hipMemcpyAsync(d_trans.ptr, trans.data(), trans.size() * 8, hipMemcpyHostToDevice, st), "copy");
    check_hip(hipMemcpyAsync(d_tr_off.ptr, tr_off.data(), (n_st + 1) * 4, hipMemcpyHostToDevice, st), "copy");
    check_hip(hipMemcpyAsync(d_atom_st.ptr, atom_st.data(), n_atoms * 4, hipMemcpyHostToDevice, st), "copy");
    check_hip(hipMemcpyAsync(d_atom_off.ptr, atom_off.data(), (n_st + 1) * 4, hipMemcpyHostToDevice, st), "copy");
    check_hip(hipMemcpyAsync(d_row_off.ptr, row_off.data(), (n_st + 1) * 4, hipMemcpyHostToDevice, st), "copy");
    // cell-list search for one large structure whose cell spans >= 3 bins of
    // width rc along every lattice direction
    if (n_st == 1 && n_atoms >= 1024) {
        CellInfo info;
        std::vector<double> tr, prev;
        cell_translations(axes, pos, n_atoms, cutoff, tr, prev, &info);
        double inv[9];
        {
            const double* a = info.axis;
            const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                               a[2] * (a[3] * a[7] - a[4] * a[6]);
            inv[0] = (a[4] * a[8] - a[5] * a[7]) / det, inv[1] = -(a[1] * a[8] - a[2] * a[7]) / det;
            inv[2] = (a[1] * a[5] - a[2] * a[4]) / det, inv[3] = -(a[3] * a[8] - a[5] * a[6]) / det;
            inv[4] = (a[0] * a[8] - a[2] * a[6]) / det, inv[5] = -(a[0] * a[5] - a[2] * a[3]) / det;
            inv[6] = (a[3] * a[7] - a[4] * a[6]) / det, inv[7] = -(a[0] * a[7] - a[1] * a[6]) / det;
            inv[8] = (a[0] * a[4] - a[1] * a[3]) / det;
        }
        bool ok = true;
        for (int a = 0; a < 3; ++a) {
            const double h = 1.0 / std::sqrt(inv[3 * a] * inv[3 * a] + inv[3 * a + 1] * inv[3 * a + 1] +
                                             inv[3 * a + 2] * inv[3 * a + 2]);
            grid.nb[a] = (int)std::floor(h / cutoff);
            grid.mexp[a] = info.mexp[a];
            if (grid.nb[a] < 3) ok = false;
        }
        if (ok) {
            binned = true;
            for (int a = 0; a < 9; ++a) grid.inv[a] = inv[a];
            d_tindex.alloc(info.tindex.size() * 4);
            check_hip(hipMemcpy(d_tindex.ptr, info.tindex.data(), info.tindex.size() * 4,
                                hipMemcpyHostToDevice), "copy");
            grid.tindex = (const int*)d_tindex.ptr;
        }
    }
    n_pairs = build_list(false, d_nb_off, d_nb_j, d_nb_t, d_nb_d, nb_off);
    max_pairs_atom = 0;
    for (int i = 0; i < n_atoms; ++i) max_pairs_atom = std::max(max_pairs_atom, nb_off[i + 1] - nb_off[i]);
}

void BatchHandle::prepare_bins() {
    Device& dev = device();
    hipStream_t st = dev.stream;
    const int nbins = grid.nb[0] * grid.nb[1] * grid.nb[2];
    d_bin.alloc((size_t)n_atoms * 4), d_wrap.alloc((size_t)n_atoms * 12);
    d_bin_start.alloc((size_t)(nbins + 1) * 4), d_bin_sorted.alloc((size_t)n_atoms * 4);
    DevBuf cnt, cur;
    cnt.alloc((size_t)nbins * 4), cur.alloc((size_t)nbins * 4);
    check_hip(hipMemsetAsync(cnt.ptr, 0, (size_t)nbins * 4, st), "memset");
    check_hip(hipMemsetAsync(cur.ptr, 0, (size_t)nbins * 4, st), "memset");
    launch_bin_assign(st, n_atoms, (const double*)d_pos.ptr, grid, (int*)d_bin.ptr, (int*)d_wrap.ptr,
                      (int*)cnt.ptr);
    check_hip(hipGetLastError(), "bin assign launch");
    std::vector<int> c(nbins), start(nbins + 1, 0);
    check_hip(hipMemcpyAsync(c.data(), cnt.ptr, (size_t)nbins * 4, hipMemcpyDeviceToHost, st), "copy");
    check_hip(hipStreamSynchronize(st), "sync");
    for (int b = 0; b < nbins; ++b) start[b + 1] = start[b] + c[b];
    check_hip(hipMemcpyAsync(d_bin_start.ptr, start.data(), (size_t)(nbins + 1) * 4,
                             hipMemcpyHostToDevice, st), "copy");
    launch_bin_scatter(st, n_atoms, (const int*)d_bin.ptr, (const int*)d_bin_start.ptr, (int*)cur.ptr,
                       (int*)d_bin_sorted.ptr);
    check_hip(hipGetLastError(), "bin scatter launch");
    check_hip(hipStreamSynchronize(st), "sync");
}

long BatchHandle::build_list(bool half, DevBuf& off, DevBuf& nj, DevBuf& nt, DevBuf& nd,
                             std::vector<int>& host_off) {
    Device& dev = device();
    hipStream_t st = dev.stream;
    const double cut2 = cutoff * cutoff;
    const double tol = 1e-10, tol2 = tol * tol;
    if (binned && !d_bin.ptr) prepare_bins();
    DevBuf cnt;
    cnt.alloc((size_t)n_atoms * 4);
    auto run = [&](int* c, const int* o, int* j, int* t, double* d) {
        if (binned)
            launch_bin_pairs(st, n_atoms, half ? 1 : 0, (const double*)d_pos.ptr, (const int*)d_bin.ptr,
                             (const int*)d_wrap.ptr, (const int*)d_bin_start.ptr,
                             (const int*)d_bin_sorted.ptr, grid, (const double*)d_trans.ptr, cut2,
                             tol2, tol, c, o, j, t, d);
        else if (half)
            launch_neighbor_half(st, n_atoms, (int*)d_atom_st.ptr, (int*)d_atom_off.ptr,
                                 (double*)d_pos.ptr, (int*)d_tr_off.ptr, (double*)d_trans.ptr, cut2,
                                 tol2, tol, c, o, j, t, d);
        else
            launch_neighbor_full(st, n_atoms, (int*)d_atom_st.ptr, (int*)d_atom_off.ptr,
                                 (double*)d_pos.ptr, (int*)d_tr_off.ptr, (double*)d_trans.ptr, cut2,
                                 tol2, c, o, j, t, d);
        check_hip(hipGetLastError(), "neighbor launch");
    };
    run((int*)cnt.ptr, nullptr, nullptr, nullptr, nullptr);
    std::vector<int> c(n_atoms);
    check_hip(hipMemcpyAsync(c.data(), cnt.ptr, (size_t)n_atoms * 4, hipMemcpyDeviceToHost, st), "copy");
    check_hip(hipStreamSynchronize(st), "sync");
    host_off.assign(n_atoms + 1, 0);
    for (int i = 0; i < n_atoms; ++i) host_off[i + 1] = host_off[i] + c[i];
    const long M = host_off[n_atoms];
    off.alloc((size_t)(n_atoms + 1) * 4);
    nj.alloc(std::max<long>(M, 1) * 4);
    nt.alloc(std::max<long>(M, 1) * 4);
    nd.alloc(std::max<long>(M, 1) * 24);
    check_hip(hipMemcpyAsync(off.ptr, host_off.data(), (size_t)(n_atoms + 1) * 4,
                             hipMemcpyHostToDevice, st), "copy");
    run(nullptr, (const int*)off.ptr, (int*)nj.ptr, (int*)nt.ptr, (double*)nd.ptr);
    check_hip(hipStreamSynchronize(st), "sync");
    return M;
}

BatchDev BatchHandle::view() const {
    BatchDev b;
    b.n_st = n_st;
    b.n_atoms_total = n_atoms;
    b.st_atom_off = (const int*)d_atom_off.ptr;
    b.st_row_off = (const int*)d_row_off.ptr;
    b.types = (const int*)d_types.ptr;
    b.atom_st = (const int*)d_atom_st.ptr;
    b.pos = (const double*)d_pos.ptr;
    b.nb_off = (const int*)d_nb_off.ptr;
    b.nb_j = (const int*)d_nb_j.ptr;
    b.nb_d = (const double*)d_nb_d.ptr;
    return b;
}

// ---------------------------------------------------------------------------
// X build for a range of structures into a row-major buffer
// ---------------------------------------------------------------------------
static constexpr size_t kLdsBytes = 160 * 1024;

XPlan plan_x(ModelHandle& m, int rows_max, bool fp64, int cover) {
    const int elem = fp64 ? 8 : 4;
    // the row tile in LDS when some shape holds it, else in X itself
    // (PLM_XROWS=global: always in X)
    const char* env = std::getenv("PLM_XROWS");
    const int g0 = env && std::string(env) == "global" ? 1 : 0;
    for (int gtile = g0; gtile <= 1; ++gtile)
    for (int CB : {64, 32, 16}) {
        const auto& ct = m.columns(CB, cover);
        const PairRec R(m.md.n_fn, m.md.n_lmh);
        const int ppw = 64 / CB;
        const int slot = ppw * R.prefix(ct.nkey_max);
        if (slot > xrows_stage_limit(fp64)) continue;
        for (int NW : {16, 8, 4}) {
            if (NW > xrows_max_waves(fp64)) continue;
            XRowsParams P{};
            P.CB = CB, P.NW = NW, P.need_max = ct.need_max, P.wz_max = ct.wz_max;
            P.rows_max = rows_max, P.ntp = m.md.n_tp, P.slot = slot, P.gtile = gtile;
            const size_t sm = xrows_smem(P, elem);
            if (sm <= kLdsBytes - 1024) {
                XPlan p{P, &ct, sm};
                return p;
            }
        }
    }
    throw std::runtime_error("column block too large for LDS");
}

// Signature-group path for the linear columns: the launch shapes whose row
// tile holds the largest structure's rows in LDS, the rule-based shape first
// (the largest TB <= 8 dividing the waves, all waves used, static grid). TB
// tasks per workgroup, A waves per task; when not even one task's tile fits,
// the tile is cut into slot parts (split 2 or 4, TB = 1). Empty when the path
// is off (PLM_XROWS=generic or global) or the model has no signature groups.
struct GroupShape {
    int TB, A, split;
    bool persistent;
};

static std::vector<GroupShape> group_shapes(ModelHandle& m, int rows_max, bool fp64) {
    std::vector<GroupShape> out;
    const char* env = std::getenv("PLM_XROWS");
    if (env && (std::string(env) == "generic" || std::string(env) == "global")) return out;
    if (!m.groups().enabled) return out;
    const int NSW = xgroup_slots(fp64), elem = 8;  // the row tile is FP64 in both precisions
    const size_t per_col = (size_t)(rows_max - 1) * elem;
    const int cap = (int)((kLdsBytes - 1024) / per_col);  // tile columns that fit, padding included
    const int waves = xgroup_max_waves(fp64);
    int split = 1;
    while (split <= 4 && NSW / split + 1 > cap) split *= 2;
    if (split > 4) return out;
    if (split > 1) {
        out.push_back({1, waves, split, false});
        out.push_back({1, waves, split, true});
        return out;
    }
    int tb_rule = std::min(8, (cap - 1) / NSW);
    while (tb_rule > 1 && waves % tb_rule) --tb_rule;
    out.push_back({tb_rule, waves / tb_rule, 1, false});
    for (int tb : {1, 2, 4, 8}) {
        if (tb * NSW + 1 > cap || waves % tb) continue;
        for (int a : {waves / tb, waves / tb / 2}) {
            if (a < 1) continue;
            for (bool pers : {false, true})
                if (!(tb == tb_rule && a == waves / tb_rule && !pers))
                    out.push_back({tb, a, 1, pers});
        }
    }
    return out;
}

// One signature-group pass (one launch per atom type) over structures [s0, s1).
static void run_group(hipStream_t st, bool fp64, ModelHandle& m, const BatchDev& bd,
                      const GroupShape& g, int s0, int s1, int row_base, int a0, long pb, int ld,
                      int rows_max, XWork& work, const PairRec& R, void* X, const double* roww) {
    const GroupTables& G = m.groups();
    const int H = kGroupSlots / xgroup_slots(fp64);
    const int nt = m.host.n_type();
    if (g.persistent) {
        work.queue.alloc((size_t)nt * 8 * sizeof(int));
        // cleared by a kernel (a memset node would not replay correctly in a graph)
        launch_fill_zero(st, (double*)work.queue.ptr, (size_t)nt * 4);
    }
    for (int t = 0; t < nt; ++t) {
        XGroupParams P{};
        P.TB = g.TB, P.A = g.A, P.split = g.split;
        P.vt0 = G.type_task0[t] * H, P.vt1 = G.type_task0[t + 1] * H;
        if (P.vt1 <= P.vt0) continue;
        P.n_blocks = (P.vt1 - P.vt0 + g.TB - 1) / g.TB;
        P.ld = ld, P.st0 = s0, P.st_end = s1, P.row_base = row_base;
        P.xcd_map = (s1 - s0) >= 16 ? 1 : 0;
        P.atom_base = a0;
        P.pair_base = pb;
        P.w_stride = std::max(m.wz_max_type, 1);
        if (g.persistent) {
            P.queue = (int*)work.queue.ptr + 8 * t;
            const long items = (long)(s1 - s0) * P.n_blocks * P.split;
            P.grid = (int)std::min<long>(items, xgroup_persistent_grid(P, rows_max, fp64));
        }
        launch_xgroup(st, fp64, G.dev, bd, P, rows_max, work.w.ptr, work.pd.ptr, R,
                      (const int*)work.tpo.ptr, X, roww);
    }
}

static bool stream_capturing(hipStream_t st) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    check_hip(hipStreamIsCapturing(st, &cs), "capture status");
    return cs != hipStreamCaptureStatusNone;
}

// The autotuner of the signature-group launch: on the first sub-batch of a
// (precision, largest-structure row bucket, structure-count bucket) class,
// every candidate shape runs the pass over the sub-batch's largest structures
// (at most 64, and at most 1 GiB of linear-column rows) into scratch rows,
// once to warm up and once timed with HIP events; the fastest is kept on the
// model for the class. Off with PLM_AUTOTUNE=0 and while the stream is being
// captured (then the rule-based shape, or the class's earlier choice, is used).
// PLM_XGROUP="TB,A,split,persistent" forces one shape.
static GroupShape choose_group(ModelHandle& m, BatchHandle& b, const BatchDev& bd, int s0, int s1,
                               int rows_max, bool fp64, int a0, long pb, XWork& work,
                               const PairRec& R, const std::vector<GroupShape>& shapes,
                               hipStream_t st) {
    // PLM_XGROUP="TB,A,split,persistent" forces a shape (the launch checks it)
    if (const char* f = std::getenv("PLM_XGROUP")) {
        GroupShape g{};
        int pers = 0;
        if (std::sscanf(f, "%d,%d,%d,%d", &g.TB, &g.A, &g.split, &pers) != 4)
            throw std::invalid_argument("PLM_XGROUP: TB,A,split,persistent");
        g.persistent = pers != 0;
        return g;
    }
    const int n_st = s1 - s0;
    int nb = 1;
    while (nb < std::min(n_st, 64)) nb *= 2;
    const auto key = std::make_tuple(fp64 ? 1 : 0, (rows_max + 95) / 96, nb);
    // a class's earlier choice, when it also holds this sub-batch's rows
    auto it = m.group_tuning.find(key);
    if (it != m.group_tuning.end())
        for (const GroupShape& g : shapes)
            if (g.TB == it->second.TB && g.A == it->second.A && g.split == it->second.split &&
                g.persistent == it->second.persistent)
                return g;
    const char* env = std::getenv("PLM_AUTOTUNE");
    if (shapes.size() == 1 || (env && std::string(env) == "0") || stream_capturing(st))
        return shapes[0];
    // tuning structures: the largest ones (batches are in ascending size order)
    const int elem = fp64 ? 8 : 4;
    const int nlin = m.host.n_linear();
    const int ldt = (nlin + 3) / 4 * 4;
    int t0 = s1;
    size_t bytes = 0;
    while (t0 > s0 && s1 - t0 < 64) {
        const size_t add = (size_t)(b.row_off[t0] - b.row_off[t0 - 1]) * ldt * elem;
        if (t0 < s1 && bytes + add > (1ull << 30)) break;
        bytes += add, --t0;
    }
    work.tune_x.alloc(bytes);
    hipEvent_t e0, e1;
    check_hip(hipEventCreate(&e0), "event");
    check_hip(hipEventCreate(&e1), "event");
    ModelHandle::GroupChoice best;
    best.ms = 1e30;
    for (size_t k = 0; k < shapes.size(); ++k) {
        const GroupShape& g = shapes[k];
        run_group(st, fp64, m, bd, g, t0, s1, b.row_off[t0], a0, pb, ldt, rows_max, work, R,
                  work.tune_x.ptr, nullptr);
        check_hip(hipEventRecord(e0, st), "event");
        run_group(st, fp64, m, bd, g, t0, s1, b.row_off[t0], a0, pb, ldt, rows_max, work, R,
                  work.tune_x.ptr, nullptr);
        check_hip(hipEventRecord(e1, st), "event");
        check_hip(hipEventSynchronize(e1), "event sync");
        float ms = 0.f;
        check_hip(hipEventElapsedTime(&ms, e0, e1), "event time");
        if (k == 0) best.rule_ms = ms;
        if (ms < best.ms) best.TB = g.TB, best.A = g.A, best.split = g.split,
                          best.persistent = g.persistent, best.ms = ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    best.candidates = (int)shapes.size();
    m.group_tuning[key] = best;
    return {best.TB, best.A, best.split, best.persistent};
}

void build_x_rows(ModelHandle& m, BatchHandle& b, int s0, int s1, bool fp64, void* X, int ld,
                  const double* roww, double* xe_sum, double* xe_sq, XWork& work, bool x_is_zero,
                  const long* poly_doff) {
    Device& dev = device();
    hipStream_t st = dev.stream;
    const int a0 = b.atom_off[s0], a1 = b.atom_off[s1];
    const int na = a1 - a0;
    const int elem = fp64 ? 8 : 4;
    const int a_stride = m.max_nlm;
    BatchDev bd = b.view();
    const int w_stride = std::max(m.wz_max_type, 1), x_stride = std::max(m.lin_max_type, 1);
    int rows_max = 0;
    for (int s = s0; s < s1; ++s) rows_max = std::max(rows_max, 7 + 3 * (b.atom_off[s + 1] - b.atom_off[s]));
    const bool use_poly = m.poly().enabled;
    const std::vector<GroupShape> shapes = group_shapes(m, rows_max, fp64);
    const GroupDev* gd = shapes.empty() ? nullptr : &m.groups().dev;
    // pair records (grouped by type pair), then the order parameters from them
    const long pb = b.nb_off[a0], pe = b.nb_off[a1];
    const PairRec R(m.md.n_fn, m.md.n_lmh);
    work.pd.alloc((size_t)std::max<long>(pe - pb, 1) * R.stride * elem);
    work.tpo.alloc((size_t)na * (m.md.n_tp + 1) * sizeof(int));
    {
        ProfScope ps("pairdata", st);
        launch_pairdata(st, fp64, m.md, bd, a0, na, pb, work.pd.ptr, R, m.md.max_l,
                        (int*)work.tpo.ptr);
    }
    check_hip(hipGetLastError(), "pairdata launch");
    work.a.alloc((size_t)na * a_stride * 2 * elem);
    {
        ProfScope ps("anlm", st);
        launch_anlm_rec(st, fp64, m.md, bd, a0, na, pb, work.pd.ptr, R, (const int*)work.tpo.ptr,
                        work.a.ptr, a_stride, m.max_noconj);
    }
    check_hip(hipGetLastError(), "anlm launch");
    work.w.alloc((size_t)na * w_stride * 2 * elem);
    work.x.alloc((size_t)na * x_stride * elem);
    {
        ProfScope ps("wprime", st);
        launch_wprime(st, fp64, m.md, bd, work.a.ptr, a_stride, a0, na, work.w.ptr, w_stride,
                      work.x.ptr, x_stride, m.max_nlm, m.w_maxf, m.v_maxf);
    }
    check_hip(hipGetLastError(), "wprime launch");
    if (gd) {
        // linear columns by signature groups: energy rows from the feature
        // values, then one launch per atom type adding its stress and force rows
        const int nlin = m.host.n_linear();
        const long rows = b.row_off[s1] - b.row_off[s0];
        if (!x_is_zero)
            check_hip(hipMemset2DAsync(X, (size_t)ld * elem, 0, (size_t)nlin * elem, (size_t)rows, st),
                      "memset");
        {
            ProfScope px("xenergy", st);
            launch_xenergy(st, fp64, *gd, bd, nlin, s0, s1 - s0, b.row_off[s0], a0, work.x.ptr,
                           x_stride, X, ld, roww, xe_sum, xe_sq);
        }
        const GroupShape g = choose_group(m, b, bd, s0, s1, rows_max, fp64, a0, pb, work, R,
                                          shapes, st);
        ProfScope pg("xgroup", st);
        run_group(st, fp64, m, bd, g, s0, s1, b.row_off[s0], a0, pb, ld, rows_max, work, R, X,
                  roww);
    } else {
        XPlan pl = plan_x(m, rows_max, fp64, use_poly ? m.host.n_linear() : m.host.n_features());
        pl.P.ld = ld;
        pl.P.st0 = s0;
        pl.P.st_end = s1;
        pl.P.row_base = b.row_off[s0];
        pl.P.xcd_map = (s1 - s0) >= 16 ? 1 : 0;
        pl.P.pair_base = pb;
        ProfScope ps("xrows", st);
        launch_xrows(st, fp64, m.md, pl.ct->cd, bd, pl.ct->bw, pl.P, work.w.ptr, w_stride,
                     work.x.ptr, x_stride, a0, work.pd.ptr, R, X, roww, xe_sum, xe_sq);
    }
    check_hip(hipGetLastError(), "xrows launch");
    if (use_poly)
        build_poly_rows(m, b, s0, s1, fp64, X, ld, roww, xe_sum, xe_sq, work, pb, R, w_stride,
                        x_stride, poly_doff);
}

size_t atom_work_bytes(ModelHandle& m, bool fp64) {
    const size_t elem = fp64 ? 8 : 4;
    const GroupTables& G = m.groups();
    const size_t grouped = G.enabled ? (size_t)(m.md.n_tp + 1) * 4 : 0;
    return elem * (2 * (size_t)m.max_nlm + 2 * (size_t)std::max(m.wz_max_type, 1) +
                   (size_t)std::max(m.lin_max_type, 1)) + grouped;
}

size_t pair_record_bytes(const ModelHandle& m, bool fp64) {
    return (size_t)PairRec(m.md.n_fn, m.md.n_lmh).stride * (fp64 ? 8 : 4);
}

}  // namespace plm

using namespace plm;

#define PLM_TRY(...)                                         \
    try {                                                    \
        __VA_ARGS__;                                         \
        return PLM_OK;                                       \
    } catch (const DeviceError& e) {                         \
        g_last_error = e.what();                             \
        return PLM_ERR_DEVICE;                               \
    } catch (const std::bad_alloc& e) {                      \
        g_last_error = "out of host memory";                 \
        return PLM_ERR_MEMORY;                               \
    } catch (const std::invalid_argument& e) {               \
        g_last_error = e.what();                             \
        return PLM_ERR_ARG;                                  \
    } catch (const std::exception& e) {                      \
        g_last_error = e.what();                             \
        return PLM_ERR_MODEL;                                \
    }

// precision -> (FP64 X build, quantised int8 Gram)
static bool fp64_features(int precision) {
    if (precision != PLM_FP64 && precision != PLM_MIXED && precision != PLM_MIXED64)
        throw std::invalid_argument("unknown precision");
    return precision != PLM_MIXED;
}
static bool quant_gram(int precision) { return precision != PLM_FP64; }

extern "C" {

const char* plm_last_error(void) { return g_last_error.c_str(); }

int plm_device_count(int* n) {
    PLM_TRY({
        int c = 0;
        if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
        *n = c;
    });
}

int plm_model_create(const plm_model_desc* d, plm_model** out) {
    PLM_TRY({
        if (!d || !out) throw std::invalid_argument("null argument");
        ModelInput in;
        in.n_type = d->n_type;
        in.gtinv = d->feature_type == 1;
        in.model_type = d->model_type;
        in.max_p = d->max_p;
        in.max_l = d->max_l;
        in.cutoff = d->cutoff;
        for (int k = 0; k < d->n_params; ++k) in.params.push_back({d->params[2 * k], d->params[2 * k + 1]});
        in.cond.assign(in.n_type, std::vector<std::vector<int>>(in.n_type));
        int tp = 0, off = 0;
        for (int i = 0; i < in.n_type; ++i)
            for (int j = i; j < in.n_type; ++j, ++tp) {
                const int c = d->cond_count[tp];
                in.cond[i][j].assign(d->cond_index + off, d->cond_index + off + c);
                off += c;
            }
        if (in.gtinv) {
            int pl = 0, pm = 0, pc = 0;
            for (int k = 0; k < d->n_lcomb; ++k) {
                const int o = d->lcomb_order[k], nt = d->lcomb_nterm[k];
                in.l_comb.emplace_back(d->lcomb_l + pl, d->lcomb_l + pl + o);
                pl += o;
                std::vector<std::vector<int>> lm(nt);
                for (int t = 0; t < nt; ++t) lm[t].assign(d->lm + pm + t * o, d->lm + pm + (t + 1) * o);
                pm += o * nt;
                in.lm_seq.push_back(lm);
                in.lm_coeffs.emplace_back(d->lm_coeff + pc, d->lm_coeff + pc + nt);
                pc += nt;
            }
        }
        auto h = std::make_unique<ModelHandle>(in);
        *out = reinterpret_cast<plm_model*>(h.release());
    });
}

void plm_model_destroy(plm_model* m) { delete reinterpret_cast<ModelHandle*>(m); }

int plm_model_n_features(const plm_model* m, int* n) {
    PLM_TRY({ *n = reinterpret_cast<const ModelHandle*>(m)->host.n_features(); });
}

int plm_model_type_sizes(const plm_model* m, int type, int* n_nlm, int* n_noconj, int* n_linear) {
    PLM_TRY({
        const auto& M = reinterpret_cast<const ModelHandle*>(m)->host;
        if (type < 0 || type >= M.n_type()) throw std::invalid_argument("type out of range");
        const auto& T = M.type(type);
        *n_nlm = M.input().gtinv ? (int)T.nlm.size() : (int)T.ntp.size();
        *n_noconj = M.input().gtinv ? (int)T.noconj.size() : (int)T.ntp.size();
        *n_linear = (int)T.features.size();
    });
}

int plm_cell_translations(const double* axis, const double* pos_c, int n, double cutoff,
                          double* trans, int max_trans, double* pos_rev, int* n_trans) {
    PLM_TRY({
        std::vector<double> tr, prev;
        cell_translations(axis, pos_c, n, cutoff, tr, prev);
        *n_trans = (int)(tr.size() / 3);
        if (trans) std::memcpy(trans, tr.data(), sizeof(double) * 3 * std::min(*n_trans, max_trans));
        if (pos_rev) std::memcpy(pos_rev, prev.data(), sizeof(double) * prev.size());
    });
}

int plm_neighbor(const double* axis, const double* pos_c, int n, double cutoff, int half,
                 int* offset, int* neigh, int* trans_id, double* d, long capacity, long* n_pairs) {
    PLM_TRY({
        std::vector<int> types(n, 0);
        BatchHandle b(nullptr, cutoff, 1, axis, pos_c, types.data(), &n);
        DevBuf off, nj, nt, nd;
        std::vector<int> host_off;
        const long M = half ? b.build_list(true, off, nj, nt, nd, host_off) : b.n_pairs;
        if (!half) host_off = b.nb_off;
        *n_pairs = M;
        if (capacity < M || !offset) return PLM_OK;
        Device& dev = device();
        hipStream_t st = dev.stream;
        const DevBuf& J = half ? nj : b.d_nb_j;
        const DevBuf& T = half ? nt : b.d_nb_t;
        const DevBuf& D = half ? nd : b.d_nb_d;
        std::memcpy(offset, host_off.data(), (size_t)(n + 1) * 4);
        if (neigh) check_hip(hipMemcpyAsync(neigh, J.ptr, M * 4, hipMemcpyDeviceToHost, st), "copy");
        if (trans_id) check_hip(hipMemcpyAsync(trans_id, T.ptr, M * 4, hipMemcpyDeviceToHost, st), "copy");
        if (d) check_hip(hipMemcpyAsync(d, D.ptr, M * 24, hipMemcpyDeviceToHost, st), "copy");
        check_hip(hipStreamSynchronize(st), "sync");
    });
}

int plm_batch_create(const plm_model* m, int n_st, const double* axes, const double* pos,
                     const int* types, const int* n_atoms, plm_batch** out) {
    PLM_TRY({
        if (!m || !out || n_st <= 0) throw std::invalid_argument("bad batch arguments");
        auto* mh = const_cast<ModelHandle*>(reinterpret_cast<const ModelHandle*>(m));
        if (!mh->built) mh->build_device();
        auto b = std::make_unique<BatchHandle>(mh, mh->host.input().cutoff, n_st, axes, pos, types,
                                               n_atoms);
        *out = reinterpret_cast<plm_batch*>(b.release());
    });
}

void plm_batch_destroy(plm_batch* b) { delete reinterpret_cast<BatchHandle*>(b); }

int plm_batch_n_rows(const plm_batch* b, long* rows) {
    PLM_TRY({ *rows = reinterpret_cast<const BatchHandle*>(b)->row_off.back(); });
}

int plm_batch_n_pairs(const plm_batch* b, long* pairs) {
    PLM_TRY({ *pairs = reinterpret_cast<const BatchHandle*>(b)->n_pairs; });
}

int plm_batch_row_order(const plm_batch* bp, long* order) {
    PLM_TRY({
        const auto* b = reinterpret_cast<const BatchHandle*>(bp);
        long r = 0;
        for (int s = 0; s < b->n_st; ++s) order[r++] = b->row_off[s];
        for (int s = 0; s < b->n_st; ++s)
            for (int k = 0; k < 6; ++k) order[r++] = b->row_off[s] + 1 + k;
        for (int s = 0; s < b->n_st; ++s) {
            const int n = b->atom_off[s + 1] - b->atom_off[s];
            for (int k = 0; k < 3 * n; ++k) order[r++] = b->row_off[s] + 7 + k;
        }
    });
}

int plm_features_x(plm_model* mp, plm_batch* bp, int precision, double* x_out) {
    PLM_TRY({
        auto* m = reinterpret_cast<ModelHandle*>(mp);
        auto* b = reinterpret_cast<BatchHandle*>(bp);
        if (!m->built) m->build_device();
        const bool fp64 = fp64_features(precision);
        const int F = m->host.n_features();
        const long R = b->row_off.back();
        const int elem = fp64 ? 8 : 4;
        Device& dev = device();
        hipStream_t st = dev.stream;
        DevBuf xb;
        XWork work;
        std::vector<double> hx((size_t)F);
        std::vector<long> order(R);
        plm_batch_row_order(bp, order.data());
        std::vector<long> inv(R);
        for (long r = 0; r < R; ++r) inv[order[r]] = r;
        // process structures in chunks bounded by ~1 GiB of X
        const size_t row_bytes = (size_t)F * elem;
        int s0 = 0;
        std::vector<char> hbuf;
        while (s0 < b->n_st) {
            int s1 = s0 + 1;
            while (s1 < b->n_st && (size_t)(b->row_off[s1 + 1] - b->row_off[s0]) * row_bytes < (1ull << 30)) ++s1;
            const long r0 = b->row_off[s0], r1 = b->row_off[s1];
            xb.alloc((size_t)(r1 - r0) * row_bytes);
            build_x_rows(*m, *b, s0, s1, fp64, xb.ptr, F, nullptr, nullptr, nullptr, work);
            hbuf.resize((size_t)(r1 - r0) * row_bytes);
            check_hip(hipMemcpyAsync(hbuf.data(), xb.ptr, hbuf.size(), hipMemcpyDeviceToHost, st), "copy");
            check_hip(hipStreamSynchronize(st), "sync");
            for (long r = r0; r < r1; ++r) {
                double* dst = x_out + (size_t)inv[r] * F;
                if (fp64) std::memcpy(dst, hbuf.data() + (size_t)(r - r0) * row_bytes, row_bytes);
                else {
                    const float* src = reinterpret_cast<const float*>(hbuf.data() + (size_t)(r - r0) * row_bytes);
                    for (int c = 0; c < F; ++c) dst[c] = src[c];
                }
            }
            s0 = s1;
        }
    });
}

int plm_normal_create(const plm_model* mp, plm_normal** out) {
    PLM_TRY({
        auto* m = const_cast<ModelHandle*>(reinterpret_cast<const ModelHandle*>(mp));
        if (!m->built) m->build_device();
        auto acc = std::make_unique<NormalHandle>();
        acc->model = m;
        const int F = m->host.n_features();
        acc->F = F;
        acc->F_pad = (F + 1 + 127) / 128 * 128;  // X columns, then the target column F
        acc->xtx.alloc((size_t)acc->F_pad * acc->F_pad * 8);
        acc->vec.alloc((size_t)acc->F_pad * 2 * 8);
        Device& dev = device();
        check_hip(hipMemsetAsync(acc->xtx.ptr, 0, acc->xtx.bytes, dev.stream), "memset");
        check_hip(hipMemsetAsync(acc->vec.ptr, 0, acc->vec.bytes, dev.stream), "memset");
        check_hip(hipStreamSynchronize(dev.stream), "sync");
        *out = reinterpret_cast<plm_normal*>(acc.release());
    });
}

void plm_normal_destroy(plm_normal* a) { delete reinterpret_cast<NormalHandle*>(a); }

int plm_normal_accumulate(plm_normal* ap, plm_batch* bp, const double* row_w, const double* y_w,
                          int precision) {
    PLM_TRY({
        auto* acc = reinterpret_cast<NormalHandle*>(ap);
        auto* b = reinterpret_cast<BatchHandle*>(bp);
        normal_accumulate(*acc, *b, row_w, y_w, fp64_features(precision), quant_gram(precision));
    });
}

int plm_normal_result(plm_normal* ap, double* xtx, double* xty, double* xe_sum, double* xe_sq_sum,
                      double* y_sq, long* n_rows) {
    PLM_TRY({
        auto* acc = reinterpret_cast<NormalHandle*>(ap);
        Device& dev = device();
        hipStream_t st = dev.stream;
        const int F = acc->F, Fp = acc->F_pad;
        check_hip(hipStreamSynchronize(st), "sync");
        // Gram of [X | y]: the upper element triangle is valid (the tiles cover
        // the upper block triangle), so mirror element-wise; column F holds
        // X^T y and element (F, F) y^T y
        if (xtx || xty) {
            std::vector<double> full((size_t)Fp * Fp);
            check_hip(hipMemcpy(full.data(), acc->xtx.ptr, full.size() * 8, hipMemcpyDeviceToHost),
                      "copy");
            if (xtx)
                for (int i = 0; i < F; ++i)
                    for (int j = 0; j < F; ++j)
                        xtx[(size_t)i * F + j] =
                            i <= j ? full[(size_t)i * Fp + j] : full[(size_t)j * Fp + i];
            if (xty)
                for (int i = 0; i < F; ++i) xty[i] = full[(size_t)i * Fp + F];
        }
        if (y_sq)
            check_hip(hipMemcpy(y_sq, (const double*)acc->xtx.ptr + (size_t)F * Fp + F, 8,
                                hipMemcpyDeviceToHost), "copy");
        std::vector<double> v((size_t)Fp * 2);
        check_hip(hipMemcpy(v.data(), acc->vec.ptr, v.size() * 8, hipMemcpyDeviceToHost), "copy");
        if (xe_sum) std::memcpy(xe_sum, v.data(), (size_t)F * 8);
        if (xe_sq_sum) std::memcpy(xe_sq_sum, v.data() + Fp, (size_t)F * 8);
        if (n_rows) *n_rows = acc->n_rows;
    });
}

int plm_model_set_coeffs(plm_model* mp, const double* coeffs, int n) {
    PLM_TRY({
        auto* m = reinterpret_cast<ModelHandle*>(mp);
        m->host.set_potential(std::vector<double>(coeffs, coeffs + n));
        if (!m->built) m->build_device();
        upload_potential(*m);
    });
}

int plm_eval(plm_model* mp, const double* axis, const double* pos_c, const int* types, int n,
             int precision, double* energy, double* forces, double* stress) {
    PLM_TRY({
        auto* m = reinterpret_cast<ModelHandle*>(mp);
        eval_single(*m, axis, pos_c, types, n, fp64_features(precision), energy, forces, stress);
    });
}

int plm_eval_noalloc(plm_model* mp, const double* axis, const double* pos_c, const int* types,
                     int n, int precision, double* energy, double* forces, double* stress) {
    PLM_TRY({
        if (!mp || !axis || !pos_c || !types) throw std::invalid_argument("null argument");
        auto* m = reinterpret_cast<ModelHandle*>(mp);
        eval_single_noalloc(*m, axis, pos_c, types, n, fp64_features(precision), energy, forces,
                            stress);
    });
}

int plm_eval_batch(plm_model* mp, plm_batch* bp, int precision, double* energies, double* forces,
                   double* stresses) {
    PLM_TRY({
        auto* m = reinterpret_cast<ModelHandle*>(mp);
        auto* b = reinterpret_cast<BatchHandle*>(bp);
        eval_batch(*m, *b, fp64_features(precision), energies, forces, stresses);
    });
}

int plm_eval_graph_create(plm_model* mp, plm_batch* bp, int precision, plm_eval_graph** out) {
    PLM_TRY({
        if (!mp || !bp || !out) throw std::invalid_argument("null argument");
        auto g = std::make_unique<EvalGraph>();
        g->m = reinterpret_cast<ModelHandle*>(mp);
        g->b = reinterpret_cast<BatchHandle*>(bp);
        g->fp64 = fp64_features(precision);
        eval_graph_capture(*g);
        *out = reinterpret_cast<plm_eval_graph*>(g.release());
    });
}

int plm_eval_graph_run(plm_eval_graph* gp, int n_steps, double* energies, double* forces,
                       double* stresses) {
    PLM_TRY({
        if (!gp) throw std::invalid_argument("null graph");
        eval_graph_run(*reinterpret_cast<EvalGraph*>(gp), n_steps, energies, forces, stresses);
    });
}

int plm_eval_graph_destroy(plm_eval_graph* gp) {
    PLM_TRY(delete reinterpret_cast<EvalGraph*>(gp));
}

int plm_ridge(int F, const double* xtx, const double* xty, const double* alphas, int n_alpha,
              double* coeffs, int* info) {
    PLM_TRY({ ridge_solve(F, xtx, xty, alphas, n_alpha, coeffs, info); });
}

int plm_normal_graph_create(plm_normal* ap, plm_batch* bp, const double* row_w, const double* y_w,
                            int precision, plm_normal_graph** out) {
    PLM_TRY({
        if (!ap || !bp || !row_w || !y_w || !out) throw std::invalid_argument("null argument");
        auto g = std::make_unique<NormalGraph>();
        g->acc = reinterpret_cast<NormalHandle*>(ap);
        g->b = reinterpret_cast<BatchHandle*>(bp);
        normal_graph_capture(*g, row_w, y_w, fp64_features(precision), quant_gram(precision));
        *out = reinterpret_cast<plm_normal_graph*>(g.release());
    });
}

int plm_normal_graph_run(plm_normal_graph* gp, int n_steps) {
    PLM_TRY({
        if (!gp) throw std::invalid_argument("null graph");
        normal_graph_run(*reinterpret_cast<NormalGraph*>(gp), n_steps);
    });
}

int plm_normal_graph_destroy(plm_normal_graph* gp) {
    PLM_TRY(delete reinterpret_cast<NormalGraph*>(gp));
}

int plm_normal_ridge(plm_normal* ap, const double* scales, const unsigned char* zero,
                     const double* alphas, int n_alpha, double* coeffs, int* info,
                     double* sq_err) {
    PLM_TRY({
        if (!ap || !scales || !alphas || !coeffs || !info) throw std::invalid_argument("null argument");
        normal_ridge(*reinterpret_cast<NormalHandle*>(ap), scales, zero, alphas, n_alpha, coeffs,
                     info, sq_err);
    });
}

int plm_normal_sq_error(plm_normal* ap, const double* scales, const unsigned char* zero,
                        const double* coeffs, int n_coef, double* sq_err) {
    PLM_TRY({
        if (!ap || !scales || !coeffs || !sq_err) throw std::invalid_argument("null argument");
        normal_sq_error(*reinterpret_cast<NormalHandle*>(ap), scales, zero, coeffs, n_coef, sq_err);
    });
}

int plm_normal_residual(plm_normal* ap, plm_batch* bp, const double* row_w, const double* y_w,
                        int precision, const double* coeffs, int n_coef, double* xtr, double* sq) {
    PLM_TRY({
        if (!ap || !bp || !row_w || !y_w || !coeffs || !xtr || !sq)
            throw std::invalid_argument("null argument");
        normal_residual(*reinterpret_cast<NormalHandle*>(ap), *reinterpret_cast<BatchHandle*>(bp),
                        row_w, y_w, fp64_features(precision), coeffs, n_coef, xtr, sq);
    });
}

int plm_normal_keep_rows(plm_normal* ap, unsigned long long max_bytes) {
    PLM_TRY({
        if (!ap) throw std::invalid_argument("null argument");
        normal_keep_rows(*reinterpret_cast<NormalHandle*>(ap), (size_t)max_bytes);
    });
}

int plm_normal_ridge_correct(plm_normal* ap, const double* scales, const unsigned char* zero,
                             const double* alphas, int n_alpha, const double* rhs, double* delta,
                             int* info, double* dad) {
    PLM_TRY({
        if (!ap || !scales || !alphas || !rhs || !delta || !info)
            throw std::invalid_argument("null argument");
        normal_ridge_correct(*reinterpret_cast<NormalHandle*>(ap), scales, zero, alphas, n_alpha,
                             rhs, delta, info, dad);
    });
}

int plm_model_tuning(const plm_model* mp, char* buf, int cap, int* len) {
    PLM_TRY({
        if (!mp) throw std::invalid_argument("null model");
        const auto* m = reinterpret_cast<const ModelHandle*>(mp);
        std::string out;
        char line[256];
        for (const auto& kv : m->group_tuning) {
            const auto& c = kv.second;
            std::snprintf(line, sizeof line,
                          "xgroup %s rows<=%d structures~%d: TB=%d A=%d split=%d %s, %.3f ms "
                          "(rule-based shape %.3f ms, %d candidates)\n",
                          std::get<0>(kv.first) ? "fp64" : "fp32", 96 * std::get<1>(kv.first),
                          std::get<2>(kv.first), c.TB, c.A, c.split,
                          c.persistent ? "persistent queue" : "static grid", c.ms, c.rule_ms,
                          c.candidates);
            out += line;
        }
        if (len) *len = (int)out.size();
        if (buf && cap > 0) {
            const int n = std::min<int>(cap - 1, (int)out.size());
            std::memcpy(buf, out.data(), n);
            buf[n] = 0;
        }
    });
}

int plm_model_feature_attrs(const plm_model* mp, int* n_linear, int* n_poly, int* radial,
                            int* gtinv, int* tcomb, int* poly, int* type_pairs) {
    PLM_TRY({
        if (!mp) throw std::invalid_argument("null model");
        const Model& M = reinterpret_cast<const ModelHandle*>(mp)->host;
        const auto& lin = M.linear_terms();
        const bool gt = M.input().gtinv;
        const int nl = (int)lin.size();
        const int np = (int)(M.comb2().size() + M.comb3().size());
        if (n_linear) *n_linear = nl;
        if (n_poly) *n_poly = np;
        for (int f = 0; f < nl; ++f) {
            if (radial) radial[f] = lin[f].n;
            if (gtinv) gtinv[f] = gt ? lin[f].lm_comb_id : -1;
            if (tcomb) {
                for (int k = 0; k < 6; ++k) tcomb[6 * f + k] = -1;
                if (gt) {
                    const auto& c = M.tp_combs()[lin[f].order][lin[f].tp_comb_id];
                    if (c.size() > 6) throw std::runtime_error("type-pair combination longer than 6");
                    for (size_t k = 0; k < c.size(); ++k) tcomb[6 * f + k] = c[k];
                } else {
                    tcomb[6 * f] = lin[f].tp_comb_id;
                }
            }
        }
        if (poly) {
            int p = 0;
            for (const auto* cs : {&M.comb2(), &M.comb3()})
                for (const auto& c : *cs) {
                    for (int k = 0; k < 3; ++k) poly[3 * p + k] = k < (int)c.size() ? c[k] : -1;
                    ++p;
                }
        }
        if (type_pairs)
            for (int a = 0; a < M.n_type(); ++a)
                for (int b = 0; b < M.n_type(); ++b) type_pairs[a * M.n_type() + b] = M.type_pair(a, b);
    });
}

int plm_solver_cg(int F, const double* xtx, const double* xty, double alpha, const double* coef0,
                  double gtol, int max_iter, double* coef, int* n_iter) {
    PLM_TRY({
        const int n = cg_solve(F, xtx, xty, alpha, coef0, gtol, max_iter, coef);
        if (n_iter) *n_iter = n;
    });
}

int plm_surrogate_n_params(int* n) { PLM_TRY({ *n = kSurrParams; }); }

int plm_surrogate_train(float* params, float* adam_m, float* adam_v, int t0, const float* x,
                        const float* y, int n, int steps, float lr, float weight_decay,
                        float* loss) {
    PLM_TRY({
        if (n < 1 || n > kSurrMaxBatch) throw std::invalid_argument("training batch size out of range");
        if (steps < 1) throw std::invalid_argument("steps must be positive");
        surrogate_run(params, adam_m, adam_v, t0, x, y, n, steps, lr, weight_decay, 0, nullptr,
                      loss);
    });
}

int plm_surrogate_grad(const float* params, const float* x, const float* y, int n, float* grad,
                       float* loss) {
    PLM_TRY({
        if (n < 1 || n > kSurrMaxBatch) throw std::invalid_argument("batch size out of range");
        surrogate_run(const_cast<float*>(params), nullptr, nullptr, 0, x, y, n, 1, 0.f, 0.f, 1,
                      grad, loss);
    });
}

int plm_surrogate_predict(const float* params, const float* x, int n, float* y) {
    PLM_TRY({
        if (n < 0) throw std::invalid_argument("negative batch size");
        surrogate_predict(params, x, n, y);
    });
}

}  // extern "C"
