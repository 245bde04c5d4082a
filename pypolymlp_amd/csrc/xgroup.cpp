// Host tables of the signature-group path for the linear columns of X.
//
// For each atom type, the local linear features are grouped by the Y keys of
// their derivative weights W'[f, .] for every type pair (the reference's
// invariants of one angular combination differ only in the radial index and
// the coefficients). Each group is cut into tasks of up to 16 features
// (slots). Per task and type pair the grouped positions run [key][slot]; a
// table maps each position to the W' nonzero of the slot's feature and key,
// so the kernel reads one key's weights of all its slots from the atom's
// feature-major W'. The W' nonzeros themselves come from
// Features::compute_features_deriv (src/cxx/src/polymlp/polymlp_features.cpp:208-248)
// as listed in ModelHandle::build_device.
#include <algorithm>
#include <map>
#include <stdexcept>
#include <vector>

#include "capi_internal.hpp"

namespace plm {

const GroupTables& ModelHandle::groups() {
    if (group_tables) return *group_tables;
    if (!built) build_device();
    auto gt = std::make_unique<GroupTables>();
    const Model& M = host;
    const int nt = M.n_type(), ntp = M.n_tp(), nlin = M.n_linear();
    constexpr int NS = kGroupSlots;
    std::vector<int> task_type, task_gcol, tt_nkey, tt_keyoff, keys, tt_woff, tt_nid;
    std::vector<std::vector<int>> type_src;  // per type: W' nonzero of each grouped position
    std::vector<int> g2l((size_t)nt * std::max(nlin, 1), -1);
    gt->type_task0.assign(1, 0);
    bool ok = nlin > 0;
    int wg_stride = 1;
    for (int t = 0; t < nt && ok; ++t) {
        const auto& wzo = host_wz_off[t];
        const auto& wy = host_wz_ylm[t];
        const auto& wn = host_wz_nid[t];
        const auto& lg = M.type(t).linear_global;
        const int L = (int)lg.size();
        for (int f = 0; f < L; ++f) g2l[(size_t)t * nlin + lg[f]] = f;
        // signature: per type pair, the sorted Y keys of the feature's nonzeros
        std::vector<std::vector<std::vector<int>>> fk(L, std::vector<std::vector<int>>(ntp));
        std::vector<std::vector<int>> fn(L, std::vector<int>(ntp, -1));
        std::map<std::vector<int>, std::vector<int>> sig_groups;
        for (int f = 0; f < L && ok; ++f) {
            std::vector<int> sig;
            for (int tp = 0; tp < ntp; ++tp) {
                auto& ks = fk[f][tp];
                for (int g = wzo[(size_t)f * ntp + tp]; g < wzo[(size_t)f * ntp + tp + 1]; ++g) {
                    ks.push_back(wy[g]);
                    if (fn[f][tp] >= 0 && fn[f][tp] != wn[g]) ok = false;  // one radial index per (feature, tp)
                    fn[f][tp] = wn[g];
                }
                std::sort(ks.begin(), ks.end());
                if (std::adjacent_find(ks.begin(), ks.end()) != ks.end()) ok = false;
                sig.push_back((int)ks.size());
                sig.insert(sig.end(), ks.begin(), ks.end());
            }
            sig_groups[sig].push_back(f);
        }
        if (!ok) break;
        std::vector<int> pos(wzo[(size_t)L * ntp], -1);
        int cur = 0;
        for (const auto& kv : sig_groups) {
            const auto& feats = kv.second;
            for (size_t k0 = 0; k0 < feats.size(); k0 += NS) {
                const int nf = (int)std::min<size_t>(NS, feats.size() - k0);
                task_type.push_back(t);
                for (int sl = 0; sl < NS; ++sl) task_gcol.push_back(sl < nf ? lg[feats[k0 + sl]] : -1);
                const int f_first = feats[k0];
                for (int tp = 0; tp < ntp; ++tp) {
                    const auto& ks = fk[f_first][tp];
                    const int nk = (int)ks.size();
                    tt_nkey.push_back(nk);
                    tt_keyoff.push_back((int)keys.size());
                    keys.insert(keys.end(), ks.begin(), ks.end());
                    tt_woff.push_back(cur);
                    for (int sl = 0; sl < NS; ++sl)
                        tt_nid.push_back(sl < nf && nk > 0 ? fn[feats[k0 + sl]][tp] : -1);
                    for (int sl = 0; sl < NS; ++sl) {
                        if (sl >= nf) continue;  // no feature: positions stay empty
                        const int f = feats[k0 + sl];
                        for (int g = wzo[(size_t)f * ntp + tp]; g < wzo[(size_t)f * ntp + tp + 1]; ++g) {
                            const int k = (int)(std::lower_bound(ks.begin(), ks.end(), wy[g]) - ks.begin());
                            pos[g] = cur + k * NS + sl;
                        }
                    }
                    cur += nk * NS;
                }
            }
        }
        std::vector<int> src(cur, -1);
        for (size_t g = 0; g < pos.size(); ++g) {
            if (pos[g] < 0) ok = false;
            else src[pos[g]] = (int)g;
        }
        type_src.push_back(std::move(src));
        wg_stride = std::max(wg_stride, cur);
        gt->type_task0.push_back((int)task_type.size());
    }
    gt->enabled = ok && ntp < 64;  // per-type-pair tables are spread over the 64 lanes
    if (gt->enabled) {
        GroupDev& d = gt->dev;
        Store& S = gt->store;
        d.n_tasks = (int)task_type.size();
        d.ntp = ntp;
        d.task_type = S.upload(task_type.data(), task_type.size() * 4);
        d.task_gcol = S.upload(task_gcol.data(), task_gcol.size() * 4);
        d.tt_nkey = S.upload(tt_nkey.data(), tt_nkey.size() * 4);
        d.tt_keyoff = S.upload(tt_keyoff.data(), tt_keyoff.size() * 4);
        d.keys = S.upload(keys.data(), keys.size() * 4);
        d.tt_woff = S.upload(tt_woff.data(), tt_woff.size() * 4);
        d.tt_nid = S.upload(tt_nid.data(), tt_nid.size() * 4);
        d.g2l = S.upload(g2l.data(), g2l.size() * 4);
        std::vector<int> src((size_t)nt * wg_stride, -1);
        for (int t = 0; t < nt; ++t)
            std::copy(type_src[t].begin(), type_src[t].end(), src.begin() + (size_t)t * wg_stride);
        d.wg_src = S.upload(src.data(), src.size() * 4);
        d.wg_stride = wg_stride;
    }
    group_tables = std::move(gt);
    return *group_tables;
}

}  // namespace plm
