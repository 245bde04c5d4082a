// Host-side model setup (see model.hpp for the reference files it follows).
#include "model.hpp"

#include <algorithm>
#include <set>

namespace plm {

namespace {

std::vector<int> erase_key(const std::vector<int>& v, size_t idx) {
    std::vector<int> k;
    k.reserve(v.size() - 1);
    for (size_t i = 0; i < v.size(); ++i)
        if (i != idx) k.push_back(v[i]);
    std::sort(k.begin(), k.end());
    return k;
}

std::vector<int> intersect_sorted(std::vector<int> a, std::vector<int> b) {
    std::sort(a.begin(), a.end());
    std::sort(b.begin(), b.end());
    std::vector<int> c;
    std::set_intersection(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(c));
    return c;
}

// All sequences of length r over 0..n-1 in lexicographic order
// (Permutenr, polymlp_model_params_gtinv.h:18-30).
void sequences(int n, int r, std::vector<int>& cur, std::vector<std::vector<int>>& out) {
    if ((int)cur.size() == r) {
        out.push_back(cur);
        return;
    }
    for (int i = 0; i < n; ++i) {
        cur.push_back(i);
        sequences(n, r, cur, out);
        cur.pop_back();
    }
}

// Index map from a sorted list of unique keys (convert_set_to_mappings,
// polymlp_structs.h:222-236).
std::map<std::vector<int>, int> index_of(const std::vector<std::vector<int>>& keys) {
    std::map<std::vector<int>, int> m;
    for (size_t i = 0; i < keys.size(); ++i) m[keys[i]] = (int)i;
    return m;
}

}  // namespace

Model::Model(const ModelInput& in) : in_(in) {
    if (in_.n_type < 1) throw std::invalid_argument("n_type must be positive");
    if (in_.params.empty()) throw std::invalid_argument("no radial parameters");
    if (in_.gtinv && in_.l_comb.empty()) throw std::invalid_argument("empty invariant tables");
    // polymlp_model_params_polynomial.cpp:40-70 (the reference exits with 8)
    if (in_.max_p > 3) throw std::invalid_argument("max_p must be smaller than or equal to 3");
    if (in_.gtinv && in_.model_type > 4) throw std::invalid_argument("Model type error");
    if (!in_.gtinv && in_.model_type > 2) throw std::invalid_argument("Model type error");
    types_.resize(in_.n_type);
    global_to_local_.resize(in_.n_type);
    set_type_pairs();
    if (in_.gtinv) {
        set_nlm();
        set_linear_terms();
    } else {
        set_ntp();
    }
    set_polynomials();
    if (in_.gtinv) set_features_gtinv();
    else set_features_pair();
    set_poly_tables();
}

// polymlp_mapping.cpp:34-74
void Model::set_type_pairs() {
    const int nt = in_.n_type, nfn = n_fn();
    type_pairs_.assign(nt, std::vector<int>(nt, 0));
    int tp = 0;
    for (int i = 0; i < nt; ++i)
        for (int j = i; j < nt; ++j) {
            type_pairs_[i][j] = type_pairs_[j][i] = tp;
            tp_types_.push_back({i, j});
            const std::vector<int>* c = nullptr;
            if (i < (int)in_.cond.size() && j < (int)in_.cond[i].size()) c = &in_.cond[i][j];
            if (!c) throw std::invalid_argument("pair_params_conditional missing a type pair");
            for (int n : *c)
                if (n < 0 || n >= nfn) throw std::invalid_argument("bad pair_params index");
            tp_to_n_.push_back(*c);
            ++tp;
        }
    n_to_tp_.assign(nfn, {});
    tpn_to_nid_.assign(tp_types_.size(), std::vector<int>(nfn, -1));
    for (int t = 0; t < (int)tp_types_.size(); ++t) {
        int id = 0;
        for (int n : tp_to_n_[t]) {
            n_to_tp_[n].push_back(t);
            tpn_to_nid_[t][n] = id++;
        }
    }
}

// polymlp_mapping.cpp:136-246, 248-280
void Model::set_nlm() {
    const int nfn = n_fn();
    const int L = in_.max_l;
    struct Lm { int l, m, ylm_key; bool conj; double cc, sign_j; };
    std::vector<Lm> lms;
    for (int l = 0; l <= L; ++l) {
        const double sign_j = (l % 2 == 0) ? 1.0 : -1.0;
        for (int m = -l; m <= l; ++m) {
            const double cc = (m % 2 == 0) ? 1.0 : -1.0;
            Lm a;
            a.l = l, a.m = m, a.cc = cc, a.sign_j = sign_j;
            if (m < 1) a.ylm_key = (l + 3) * l / 2 + m, a.conj = false;
            else a.ylm_key = (l + 3) * l / 2 - m, a.conj = true;
            lms.push_back(a);
        }
    }
    int gid = 0;
    for (int n = 0; n < nfn; ++n) {
        const auto& tpn = n_to_tp_[n];
        for (int lm = 0; lm < (int)lms.size(); ++lm) {
            const auto& a = lms[lm];
            const int conj_sub = 2 * a.m * (int)tpn.size();
            for (int tp : tpn) {
                NlmAttr x{};
                x.n = n, x.n_id = tpn_to_nid_[tp][n], x.l = a.l, x.m = a.m, x.lm = lm;
                x.ylm_key = a.ylm_key, x.tp = tp, x.conj = a.conj, x.cc = a.cc;
                x.sign_j = a.sign_j, x.global_id = gid, x.global_conj_id = gid - conj_sub;
                nlm_global_.push_back(x);
                nlm_to_global_[{n, lm, tp}] = gid;
                ++gid;
            }
        }
    }
    // local ids per type (set_nlmtp_local_attrs)
    for (const auto& g : nlm_global_) {
        const int t1 = tp_types_[g.tp][0], t2 = tp_types_[g.tp][1];
        auto& T1 = types_[t1];
        auto& T2 = types_[t2];
        const int l1 = (int)T1.nlm.size(), l2 = (int)T2.nlm.size();
        const int nc1 = (int)T1.noconj.size(), nc2 = (int)T2.noconj.size();
        NlmAttr a = g;
        a.ilocal = l1, a.ilocal_noconj = nc1, a.jlocal_noconj = nc2;
        T1.nlm.push_back(a);
        if (!g.conj) T1.noconj.push_back(l1);
        global_to_local_[t1][g.global_id] = l1;
        if (t1 != t2) {
            NlmAttr b = g;
            b.ilocal = l2, b.ilocal_noconj = nc2, b.jlocal_noconj = nc1;
            T2.nlm.push_back(b);
            if (!g.conj) T2.noconj.push_back(l2);
            global_to_local_[t2][g.global_id] = l2;
        }
    }
    for (int t = 0; t < in_.n_type; ++t) {
        auto& T = types_[t];
        T.local_to_noconj.assign(T.nlm.size(), -1);
        for (size_t k = 0; k < T.noconj.size(); ++k) T.local_to_noconj[T.noconj[k]] = (int)k;
        for (auto& a : T.nlm) a.ilocal_conj = global_to_local_[t].at(a.global_conj_id);
    }
}

// polymlp_mapping.cpp:76-133 (pair features)
void Model::set_ntp() {
    int gid = 0;
    std::vector<std::array<int, 3>> global;  // (n, n_id, tp)
    for (int tp = 0; tp < (int)tp_types_.size(); ++tp)
        for (int n : tp_to_n_[tp]) global.push_back({n, tpn_to_nid_[tp][n], tp}), ++gid;
    for (int g = 0; g < (int)global.size(); ++g) {
        const int tp = global[g][2];
        const int t1 = tp_types_[tp][0], t2 = tp_types_[tp][1];
        const int l1 = (int)types_[t1].ntp.size(), l2 = (int)types_[t2].ntp.size();
        types_[t1].ntp.push_back({global[g][0], global[g][1], tp, l2});
        types_[t1].linear_global.push_back(g);
        if (t1 != t2) {
            types_[t2].ntp.push_back({global[g][0], global[g][1], tp, l1});
            types_[t2].linear_global.push_back(g);
        }
    }
    n_linear_ = (int)global.size();
    // model_type 2 uses every pair feature in the polynomial
    for (int g = 0; g < (int)global.size(); ++g) {
        LinearTerm lt{global[g][0], 0, global[g][2], 1, {}};
        const int tp = global[g][2];
        for (int t = 0; t < in_.n_type; ++t)
            if (tp_types_[tp][0] == t || tp_types_[tp][1] == t) lt.type1.push_back(t);
        linear_.push_back(lt);
    }
}

// polymlp_model_params_gtinv.cpp:53-151
void Model::set_linear_terms() {
    const int order_max = (int)in_.l_comb.back().size();
    const int ntp = (int)tp_types_.size();
    auto valid_types = [&](const std::vector<int>& comb) {
        std::vector<int> out;
        for (int t = 0; t < in_.n_type; ++t) {
            bool ok = true;
            for (int tp : comb)
                if (tp_types_[tp][0] != t && tp_types_[tp][1] != t) { ok = false; break; }
            if (ok) out.push_back(t);
        }
        return out;
    };
    tp_combs_.assign(order_max + 1, {});
    std::vector<std::vector<std::vector<int>>> nonzero_n(order_max + 1);
    for (int order = 1; order <= order_max; ++order) {
        std::vector<std::vector<int>> all;
        std::vector<int> cur;
        sequences(ntp, order, cur, all);
        for (auto& c : all)
            if (!valid_types(c).empty()) tp_combs_[order].push_back(c);
        for (auto& c : tp_combs_[order]) {
            std::vector<int> inter = tp_to_n_[c[0]];
            for (size_t k = 1; k < c.size(); ++k) inter = intersect_sorted(inter, tp_to_n_[c[k]]);
            nonzero_n[order].push_back(inter);
        }
    }
    std::vector<std::vector<LinearTerm>> per_n(n_fn());
    for (int id = 0; id < (int)in_.l_comb.size(); ++id) {
        const auto& lc = in_.l_comb[id];
        const int order = (int)lc.size();
        std::set<std::vector<std::pair<int, int>>> uniq;
        for (const auto& c : tp_combs_[order]) {
            std::vector<std::pair<int, int>> v;
            for (int k = 0; k < order; ++k) v.push_back({lc[k], c[k]});
            std::sort(v.begin(), v.end());
            uniq.insert(v);
        }
        for (const auto& v : uniq) {
            std::vector<int> comb;
            for (const auto& p : v) comb.push_back(p.second);
            const auto it = std::find(tp_combs_[order].begin(), tp_combs_[order].end(), comb);
            const int cid = (int)(it - tp_combs_[order].begin());
            const auto t1a = valid_types(comb);
            for (int n : nonzero_n[order][cid]) per_n[n].push_back({n, id, cid, order, t1a});
        }
    }
    for (auto& v : per_n)
        for (auto& lt : v) linear_.push_back(lt);
    n_linear_ = (int)linear_.size();
}

// polymlp_model_params_polynomial.cpp:12-274
void Model::set_polynomials() {
    const int nt = in_.n_type;
    comb1_idx_.assign(nt, {});
    comb2_idx_.assign(nt, {});
    comb3_idx_.assign(nt, {});
    if (in_.model_type == 2) {
        for (int i = 0; i < n_linear_; ++i) poly_indices_.push_back(i);
    } else if (in_.model_type > 2) {
        const int max_order = (in_.model_type == 3) ? 1 : 2;
        for (int i = 0; i < n_linear_; ++i)
            if (linear_[i].order <= max_order) poly_indices_.push_back(i);
    }
    for (int i = 0; i < n_linear_; ++i)
        for (int t : linear_[i].type1) comb1_idx_[t].push_back(i);
    auto inter = [&](std::initializer_list<int> ids) {
        std::vector<int> r;
        bool first = true;
        for (int id : ids) {
            if (first) r = linear_[id].type1, first = false;
            else r = intersect_sorted(r, linear_[id].type1);
        }
        return r;
    };
    const auto& ia = poly_indices_;
    if (in_.model_type > 1 && in_.max_p > 1) {
        int ic = 0;
        for (size_t i1 = 0; i1 < ia.size(); ++i1)
            for (size_t i2 = 0; i2 <= i1; ++i2) {
                const auto r = inter({ia[i1], ia[i2]});
                if (r.empty()) continue;
                comb2_.push_back({ia[i2], ia[i1]});
                for (int t : r) comb2_idx_[t].push_back(ic);
                ++ic;
            }
    }
    if (in_.model_type > 1 && in_.max_p > 2) {
        int ic = 0;
        for (size_t i1 = 0; i1 < ia.size(); ++i1)
            for (size_t i2 = 0; i2 <= i1; ++i2)
                for (size_t i3 = 0; i3 <= i2; ++i3) {
                    const auto r = inter({ia[i1], ia[i2], ia[i3]});
                    if (r.empty()) continue;
                    comb3_.push_back({ia[i3], ia[i2], ia[i1]});
                    for (int t : r) comb3_idx_[t].push_back(ic);
                    ++ic;
                }
    }
    n_variables_ = n_linear_ + (int)comb2_.size() + (int)comb3_.size();
}

// polymlp_features_utils.cpp:48-132, polymlp_features.cpp:80-157,
// polymlp_features_mapping_utils.cpp:24-169
void Model::set_features_gtinv() {
    for (int t = 0; t < in_.n_type; ++t) {
        auto& T = types_[t];
        std::vector<std::vector<SingleTerm>> reps;
        for (int li = 0; li < n_linear_; ++li) {
            const auto& lt = linear_[li];
            if (std::find(lt.type1.begin(), lt.type1.end(), t) == lt.type1.end()) continue;
            const auto& comb = tp_combs_[lt.order][lt.tp_comb_id];
            const auto& lm_list = in_.lm_seq[lt.lm_comb_id];
            const auto& cf = in_.lm_coeffs[lt.lm_comb_id];
            std::vector<SingleTerm> feat, rep;
            for (size_t i = 0; i < lm_list.size(); ++i) {
                std::vector<int> g1;
                for (size_t k = 0; k < lm_list[i].size(); ++k)
                    g1.push_back(nlm_to_global_.at({lt.n, lm_list[i][k], comb[k]}));
                std::sort(g1.begin(), g1.end());
                std::vector<int> g2;
                for (int g : g1) g2.push_back(nlm_global_[g].global_conj_id);
                std::sort(g2.begin(), g2.end());
                const auto& gr = std::min(g1, g2);
                std::vector<int> l1, lr;
                for (int g : g1) l1.push_back(global_to_local_[t].at(g));
                for (int g : gr) lr.push_back(global_to_local_[t].at(g));
                feat.push_back({cf[i], l1});
                rep.push_back({cf[i], lr});
            }
            T.features.push_back(feat);
            reps.push_back(rep);
            T.linear_global.push_back(li);
        }
        // value products (set_mappings_efficient) and per-feature maps
        std::set<std::vector<int>> uniq;
        for (const auto& f : reps)
            for (const auto& s : f) uniq.insert(s.ids);
        T.prod.assign(uniq.begin(), uniq.end());
        const auto pidx = index_of(T.prod);
        for (const auto& f : reps) {
            std::map<int, double> acc;
            for (const auto& s : f) acc[pidx.at(s.ids)] += s.coeff;
            T.feat_val.emplace_back(acc.begin(), acc.end());
        }
        // derivative products (set_deriv_mappings, eliminate_conj = false)
        std::set<std::vector<int>> uniqd;
        for (const auto& f : T.features)
            for (const auto& s : f)
                for (size_t i = 0; i < s.ids.size(); ++i) uniqd.insert(erase_key(s.ids, i));
        T.prod_deriv.assign(uniqd.begin(), uniqd.end());
        const auto didx = index_of(T.prod_deriv);
        for (const auto& f : T.features) {
            std::map<std::pair<int, int>, double> acc;  // (head, prod) sorted
            for (const auto& s : f)
                for (size_t i = 0; i < s.ids.size(); ++i)
                    acc[{s.ids[i], didx.at(erase_key(s.ids, i))}] += s.coeff;
            std::vector<DerivTerm> d;
            for (const auto& kv : acc) d.push_back({kv.second, kv.first.first, kv.first.second});
            T.feat_deriv.push_back(d);
        }
    }
}

// polymlp_features_utils.cpp:24-45 and set_mappings_standard (pair)
void Model::set_features_pair() {
    for (int t = 0; t < in_.n_type; ++t) {
        auto& T = types_[t];
        for (int k = 0; k < (int)T.ntp.size(); ++k) {
            T.features.push_back({{1.0, {k}}});
            T.prod.push_back({k});
            T.feat_val.push_back({{k, 1.0}});
            T.feat_deriv.push_back({{1.0, k, 0}});
        }
        T.prod_deriv = {{}};
    }
}

// polymlp_features_polynomial.cpp:22-73
void Model::set_poly_tables() {
    for (int t = 0; t < in_.n_type; ++t) {
        auto& T = types_[t];
        std::map<int, int> map1;
        int local = 0;
        for (int i : comb1_idx_[t]) {
            T.poly.push_back({i, {local}});
            map1[i] = local++;
        }
        for (int i : comb2_idx_[t])
            T.poly.push_back({n_linear_ + i, {map1.at(comb2_[i][0]), map1.at(comb2_[i][1])}});
        const int b3 = n_linear_ + (int)comb2_.size();
        for (int i : comb3_idx_[t])
            T.poly.push_back(
                {b3 + i, {map1.at(comb3_[i][0]), map1.at(comb3_[i][1]), map1.at(comb3_[i][2])}});
        std::set<std::vector<int>> uniq;
        for (const auto& p : T.poly)
            for (size_t i = 0; i < p.local_ids.size(); ++i) uniq.insert(erase_key(p.local_ids, i));
        T.prod_features.assign(uniq.begin(), uniq.end());
    }
}

// polymlp_potential.cpp:25-156 with the factor 2 of polymlp_eval.cpp:12-20
void Model::set_potential(const std::vector<double>& coeffs) {
    if ((int)coeffs.size() != n_variables_)
        throw std::invalid_argument("number of coefficients does not match the model");
    const bool elim_conj = in_.gtinv;
    for (int t = 0; t < in_.n_type; ++t) {
        auto& T = types_[t];
        const auto didx = index_of(T.prod_deriv);
        const auto fidx = index_of(T.prod_features);
        std::map<std::array<int, 3>, std::array<double, 2>> acc;
        for (const auto& term : T.poly) {
            int n_prods = 0;
            for (int id : term.local_ids) n_prods += (int)T.features[id][0].ids.size();
            const double c = 2.0 * coeffs[term.global_id];
            for (size_t i = 0; i < term.local_ids.size(); ++i) {
                const int head_feature = term.local_ids[i];
                const int pf = fidx.at(erase_key(term.local_ids, i));
                for (const auto& s : T.features[head_feature]) {
                    const double cf = c * s.coeff;
                    const double ce = cf / double(n_prods);
                    for (size_t j = 0; j < s.ids.size(); ++j) {
                        const int head = s.ids[j];
                        if (elim_conj && T.nlm[head].conj) continue;
                        const int pid = didx.at(erase_key(s.ids, j));
                        auto& v = acc[{head, pid, pf}];
                        v[0] += ce;
                        v[1] += cf;
                    }
                }
            }
        }
        const int n_heads = in_.gtinv ? (int)T.noconj.size() : (int)T.ntp.size();
        T.pot.assign(n_heads, {});
        for (const auto& kv : acc) {
            const int head = kv.first[0];
            const int slot = in_.gtinv ? T.local_to_noconj[head] : head;
            T.pot[slot].push_back({kv.second[0], kv.second[1], kv.first[1], kv.first[2]});
        }
        for (auto& v : T.pot)
            std::sort(v.begin(), v.end(), [](const PotTerm& a, const PotTerm& b) {
                if (a.prod != b.prod) return a.prod < b.prod;
                return a.prod_features < b.prod_features;
            });
    }
    has_pot_ = true;
}

}  // namespace plm
