// Host-side model setup for the polymlp hot path.
//
// Builds, from the params dict contents, every index table the device kernels
// read: type pairs and radial parameters, the (n, l, m, type-pair) order
// parameter layout, the linear invariant terms and polynomial columns of X,
// the per-type invariant term lists (values and derivatives) and the folded
// potential used by evaluation. The enumeration orders follow the reference
// exactly because they define the column order of X:
//   Mapping            src/cxx/src/polymlp/polymlp_mapping.cpp:11-280
//   uniq_gtinv_type    src/cxx/src/polymlp/polymlp_model_params_gtinv.cpp:92-151
//   ModelParamsPoly    src/cxx/src/polymlp/polymlp_model_params_polynomial.cpp:12-287
//   Features maps      src/cxx/src/polymlp/polymlp_features.cpp:27-157,
//                      polymlp_features_utils.cpp:48-132,
//                      polymlp_features_mapping_utils.cpp:24-169
//   FeaturesPoly       src/cxx/src/polymlp/polymlp_features_polynomial.cpp:22-83
//   Potential          src/cxx/src/polymlp/polymlp_potential.cpp:25-156
#pragma once

#include <array>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace plm {

struct ModelInput {
    int n_type = 1;
    bool gtinv = true;  // feature_type: gtinv (true) or pair (false)
    int model_type = 2;
    int max_p = 2;
    int max_l = 0;
    double cutoff = 6.0;
    std::vector<std::array<double, 2>> params;          // (beta, mu) per radial function
    std::vector<std::vector<std::vector<int>>> cond;    // cond[i][j] (i <= j): active n
    // invariant tables (already screened, lm = l*l + l + m)
    std::vector<std::vector<int>> l_comb;
    std::vector<std::vector<std::vector<int>>> lm_seq;
    std::vector<std::vector<double>> lm_coeffs;
};

// One order-parameter component a_{n,l,m,tp} of an atom of a given type.
struct NlmAttr {
    int n, n_id, l, m, lm, ylm_key, tp;
    bool conj;
    double cc, sign_j;
    int global_id, global_conj_id;
    int ilocal, ilocal_noconj, jlocal_noconj, ilocal_conj;
};

struct LinearTerm {
    int n, lm_comb_id, tp_comb_id, order;
    std::vector<int> type1;
};

struct PolyTerm {
    int global_id;
    std::vector<int> local_ids;  // indices into the per-type linear feature vector
};

struct SingleTerm {
    double coeff;
    std::vector<int> ids;  // local nlmtp ids (sorted by global id)
};

struct DerivTerm {
    double coeff;
    int head, prod;  // local nlmtp id of the head, id of the product of the rest
};

struct PotTerm {
    double coeff_e, coeff_f;
    int prod, prod_features;
};

struct TypeTables {
    std::vector<NlmAttr> nlm;         // all local components (incl. conjugates)
    std::vector<int> noconj;          // local ids of non-conjugate components
    std::vector<int> local_to_noconj; // local id -> noconj index (-1 for conj)
    // pair features: local ntp list (n_id, tp) in local order
    std::vector<std::array<int, 4>> ntp;  // (n, n_id, tp, jlocal)
    std::vector<int> linear_global;   // global linear feature id of each local feature
    std::vector<std::vector<SingleTerm>> features;       // deriv path (non-representative)
    std::vector<std::vector<int>> prod;                  // value products (representatives)
    std::vector<std::vector<std::pair<int, double>>> feat_val;  // feature -> (prod, coeff)
    std::vector<std::vector<int>> prod_deriv;            // products of all-but-head
    std::vector<std::vector<DerivTerm>> feat_deriv;      // feature -> deriv terms (sorted)
    std::vector<PolyTerm> poly;                          // polynomial columns of this type
    std::vector<std::vector<int>> prod_features;         // products of features (poly)
    std::vector<std::vector<PotTerm>> pot;               // per noconj head (eval)
};

class Model {
  public:
    explicit Model(const ModelInput& in);

    const ModelInput& input() const { return in_; }
    int n_type() const { return in_.n_type; }
    int n_fn() const { return (int)in_.params.size(); }
    int n_tp() const { return (int)tp_types_.size(); }
    int n_features() const { return n_variables_; }
    int n_linear() const { return n_linear_; }
    int n_lm_half() const { return (in_.max_l + 1) * (in_.max_l + 2) / 2; }
    int type_pair(int t1, int t2) const { return type_pairs_[t1][t2]; }
    const std::vector<std::array<int, 2>>& tp_types() const { return tp_types_; }
    const std::vector<std::vector<int>>& tp_to_n() const { return tp_to_n_; }
    const TypeTables& type(int t) const { return types_[t]; }
    const std::vector<LinearTerm>& linear_terms() const { return linear_; }
    const std::vector<std::vector<int>>& comb2() const { return comb2_; }
    const std::vector<std::vector<int>>& comb3() const { return comb3_; }
    const std::vector<std::vector<std::vector<int>>>& tp_combs() const { return tp_combs_; }
    const std::vector<std::vector<int>>& type_pairs() const { return type_pairs_; }

    // Fold fitted coefficients (as passed to PolymlpEval, i.e. coeffs/scales;
    // the factor 2 of polymlp_eval.cpp:17 is applied here) into per-head terms.
    void set_potential(const std::vector<double>& coeffs);
    bool has_potential() const { return has_pot_; }

  private:
    ModelInput in_;
    std::vector<std::vector<int>> type_pairs_;
    std::vector<std::array<int, 2>> tp_types_;
    std::vector<std::vector<int>> tp_to_n_;
    std::vector<std::vector<int>> n_to_tp_;
    std::vector<std::vector<int>> tpn_to_nid_;
    std::vector<std::vector<std::vector<int>>> tp_combs_;  // [order][i] -> tp list
    std::vector<NlmAttr> nlm_global_;
    std::map<std::array<int, 3>, int> nlm_to_global_;  // (n, lm, tp) -> global id
    std::vector<std::map<int, int>> global_to_local_;   // per type
    std::vector<LinearTerm> linear_;
    std::vector<int> poly_indices_;
    std::vector<std::vector<int>> comb2_, comb3_;
    std::vector<std::vector<int>> comb1_idx_, comb2_idx_, comb3_idx_;
    int n_linear_ = 0, n_variables_ = 0;
    std::vector<TypeTables> types_;
    bool has_pot_ = false;

    void set_type_pairs();
    void set_nlm();
    void set_ntp();
    void set_linear_terms();
    void set_polynomials();
    void set_features_gtinv();
    void set_features_pair();
    void set_poly_tables();
};

}  // namespace plm
