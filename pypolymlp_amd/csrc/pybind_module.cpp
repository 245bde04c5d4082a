// pybind11 module mirroring the hot-path classes of pypolymlp's libmlpcpp
// (src/pypolymlp/cxx/src/python/pybind11_mlp.cpp:10-268) on top of the C-ABI
// of the HIP library (include/polymlp_amd.h). Same class names, constructor
// signatures, getters and array layouts as the reference.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "polymlp_amd.h"

namespace py = pybind11;
using vector1i = std::vector<int>;
using vector2i = std::vector<vector1i>;
using vector1d = std::vector<double>;
using vector2d = std::vector<vector1d>;
using vector3d = std::vector<vector2d>;

namespace {

void check(int rc) {
    if (rc != PLM_OK) throw std::runtime_error(plm_last_error());
}

// params dict -> plm_model (convert_params_dict_to_feature_params, py_params.cpp:11-60)
struct ModelOwner {
    plm_model* m = nullptr;
    explicit ModelOwner(const py::dict& params) {
        const int n_type = params["n_type"].cast<int>();
        const py::dict model = params["model"].cast<py::dict>();
        std::vector<double> pp;
        for (const auto& p : model["pair_params"].cast<vector2d>()) pp.push_back(p[0]), pp.push_back(p[1]);
        const py::dict cond = model["pair_params_conditional"].cast<py::dict>();
        vector1i cc, ci;
        for (int i = 0; i < n_type; ++i)
            for (int j = i; j < n_type; ++j) {
                const auto v = cond[py::make_tuple(i, j)].cast<vector1i>();
                cc.push_back((int)v.size());
                ci.insert(ci.end(), v.begin(), v.end());
            }
        if (ci.empty()) ci.push_back(0);
        const py::dict g = model["gtinv"].cast<py::dict>();
        const auto lm_seq = g["lm_seq"].cast<std::vector<vector2i>>();
        const auto l_comb = g["l_comb"].cast<vector2i>();
        const auto coeffs = g["lm_coeffs"].cast<vector2d>();
        vector1i ord, lf, nt, lm;
        vector1d cf;
        for (size_t k = 0; k < l_comb.size(); ++k) {
            ord.push_back((int)l_comb[k].size());
            lf.insert(lf.end(), l_comb[k].begin(), l_comb[k].end());
            nt.push_back((int)coeffs[k].size());
            for (const auto& row : lm_seq[k]) lm.insert(lm.end(), row.begin(), row.end());
            cf.insert(cf.end(), coeffs[k].begin(), coeffs[k].end());
        }
        plm_model_desc d{};
        d.n_type = n_type;
        d.feature_type = model["feature_type"].cast<std::string>() == "gtinv" ? 1 : 0;
        d.model_type = model["model_type"].cast<int>();
        d.max_p = model["max_p"].cast<int>();
        d.max_l = model["max_l"].cast<int>();
        d.cutoff = model["cutoff"].cast<double>();
        d.n_params = (int)pp.size() / 2;
        d.params = pp.data();
        d.cond_count = cc.data();
        d.cond_index = ci.data();
        d.n_lcomb = (int)l_comb.size();
        d.lcomb_order = ord.data();
        d.lcomb_l = lf.data();
        d.lcomb_nterm = nt.data();
        d.lm = lm.data();
        d.lm_coeff = cf.data();
        check(plm_model_create(&d, &m));
    }
    ~ModelOwner() { plm_model_destroy(m); }
    int n_features() const {
        int n = 0;
        check(plm_model_n_features(m, &n));
        return n;
    }
};

struct Packed {
    std::vector<double> axes, pos;
    std::vector<int> types, n_atoms;
};

Packed pack(const vector3d& axis, const vector3d& positions_c, const vector2i& types) {
    Packed p;
    const size_t n_st = axis.size();
    if (positions_c.size() != n_st || types.size() != n_st)
        throw std::invalid_argument("inconsistent structure lists");
    for (size_t s = 0; s < n_st; ++s) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) p.axes.push_back(axis[s][i][j]);
        const int n = (int)types[s].size();
        for (int a = 0; a < 3; ++a) {
            if ((int)positions_c[s][a].size() != n) throw std::invalid_argument("positions shape");
            p.pos.insert(p.pos.end(), positions_c[s][a].begin(), positions_c[s][a].end());
        }
        p.types.insert(p.types.end(), types[s].begin(), types[s].end());
        p.n_atoms.push_back(n);
    }
    return p;
}

// PotentialModel (py_model.cpp:10-111)
class PotentialModel {
  public:
    PotentialModel(const py::dict& params, const vector3d& axis, const vector3d& positions_c,
                   const vector2i& types, const vector1i& n_st_dataset,
                   const std::vector<bool>& force_dataset, const vector1i& n_atoms_all) {
        ModelOwner mo(params);
        const int F = mo.n_features();
        Packed p = pack(axis, positions_c, types);
        const int n_st = (int)p.n_atoms.size();
        plm_batch* b = nullptr;
        check(plm_batch_create(mo.m, n_st, p.axes.data(), p.pos.data(), p.types.data(),
                               p.n_atoms.data(), &b));
        long rows_all = 0;
        plm_batch_n_rows(b, &rows_all);
        std::vector<double> xall((size_t)rows_all * F);
        const int rc = plm_features_x(mo.m, b, PLM_FP64, xall.data());
        plm_batch_destroy(b);
        check(rc);
        // row layout with per-dataset force flags (set_index, py_model.cpp:58-106)
        const int n_datasets = (int)n_st_dataset.size();
        std::vector<bool> force_st(n_st, false);
        xs_begin_ = xf_begin_ = vector1i(n_datasets, -1);
        n_data_ = vector1i(3, 0);
        n_data_[0] = n_st;
        std::vector<int> srow(n_st, -1), frow(n_st, -1);
        int id = 0, istress = n_st;
        for (int d = 0; d < n_datasets; ++d) {
            if (force_dataset[d]) xs_begin_[d] = istress, n_data_[2] += n_st_dataset[d] * 6;
            for (int j = 0; j < n_st_dataset[d]; ++j, ++id)
                if (force_dataset[d]) srow[id] = istress, istress += 6;
        }
        id = 0;
        int iforce = istress;
        for (int d = 0; d < n_datasets; ++d) {
            if (force_dataset[d]) xf_begin_[d] = iforce;
            for (int j = 0; j < n_st_dataset[d]; ++j, ++id)
                if (force_dataset[d]) {
                    force_st[id] = true, frow[id] = iforce;
                    iforce += 3 * n_atoms_all[id];
                    n_data_[1] += 3 * n_atoms_all[id];
                }
        }
        const long total = (long)n_data_[0] + n_data_[1] + n_data_[2];
        // reference layout of xall: energies, stresses (6 per st), forces (3N per st)
        x_ = py::array_t<double, py::array::f_style>({(py::ssize_t)total, (py::ssize_t)F});
        auto X = x_.mutable_unchecked<2>();
        std::vector<long> st_s(n_st), st_f(n_st);
        long rs = n_st, rf = n_st + 6L * n_st;
        for (int s = 0; s < n_st; ++s) st_s[s] = rs, rs += 6, st_f[s] = rf, rf += 3L * p.n_atoms[s];
        auto copy_row = [&](long dst, long src) {
            for (int c = 0; c < F; ++c) X(dst, c) = xall[(size_t)src * F + c];
        };
        for (int s = 0; s < n_st; ++s) copy_row(s, s);
        for (int s = 0; s < n_st; ++s) {
            if (!force_st[s]) continue;
            for (int k = 0; k < 6; ++k) copy_row(srow[s] + k, st_s[s] + k);
            for (int k = 0; k < 3 * p.n_atoms[s]; ++k) copy_row(frow[s] + k, st_f[s] + k);
        }
    }
    py::array_t<double, py::array::f_style>& get_x() { return x_; }
    const vector1i& get_fbegin() const { return xf_begin_; }
    const vector1i& get_sbegin() const { return xs_begin_; }
    const vector1i& get_n_data() const { return n_data_; }

  private:
    py::array_t<double, py::array::f_style> x_;
    vector1i xf_begin_, xs_begin_, n_data_;
};

// PotentialHybridModel (py_hybrid_model.cpp:11-199): X of several models side
// by side; a model with type_full false sees only the atoms whose type is in
// its type_indices (renumbered by position there), and its force rows of the
// other atoms are zero. Row layout as PotentialModel.
class PotentialHybridModel {
  public:
    PotentialHybridModel(const std::vector<py::dict>& params_array, const vector3d& axis,
                         const vector3d& positions_c, const vector2i& types,
                         const vector1i& n_st_dataset, const std::vector<bool>& force_dataset,
                         const vector1i& n_atoms_all) {
        const int n_st = (int)axis.size();
        if ((int)positions_c.size() != n_st || (int)types.size() != n_st)
            throw std::invalid_argument("axis, positions and types differ in length");
        const int n_datasets = (int)n_st_dataset.size();
        std::vector<bool> force_st(n_st, false);
        xs_begin_ = xf_begin_ = vector1i(n_datasets, -1);
        n_data_ = vector1i(3, 0);
        n_data_[0] = n_st;
        std::vector<int> srow(n_st, -1), frow(n_st, -1);
        int id = 0, istress = n_st;
        for (int d = 0; d < n_datasets; ++d) {
            if (force_dataset[d]) xs_begin_[d] = istress, n_data_[2] += n_st_dataset[d] * 6;
            for (int j = 0; j < n_st_dataset[d]; ++j, ++id)
                if (force_dataset[d]) srow[id] = istress, istress += 6;
        }
        id = 0;
        int iforce = istress;
        for (int d = 0; d < n_datasets; ++d) {
            if (force_dataset[d]) xf_begin_[d] = iforce;
            for (int j = 0; j < n_st_dataset[d]; ++j, ++id)
                if (force_dataset[d]) {
                    force_st[id] = true, frow[id] = iforce;
                    iforce += 3 * n_atoms_all[id];
                    n_data_[1] += 3 * n_atoms_all[id];
                }
        }
        std::vector<std::unique_ptr<ModelOwner>> models;
        int F = 0;
        for (const auto& pd : params_array) {
            models.push_back(std::make_unique<ModelOwner>(pd));
            F += models.back()->n_features();
            cum_.push_back(F);
        }
        const long total = (long)n_data_[0] + n_data_[1] + n_data_[2];
        x_ = py::array_t<double, py::array::f_style>({(py::ssize_t)total, (py::ssize_t)F});
        auto X = x_.mutable_unchecked<2>();
        for (long r = 0; r < total; ++r)
            for (int c = 0; c < F; ++c) X(r, c) = 0.0;
        for (size_t n = 0; n < models.size(); ++n) {
            const py::dict& pd = params_array[n];
            const bool full = !pd.contains("type_full") || pd["type_full"].cast<bool>();
            const vector1i idx = pd.contains("type_indices") ? pd["type_indices"].cast<vector1i>()
                                                             : vector1i();
            const int c0 = n ? cum_[n - 1] : 0;
            const int Fn = models[n]->n_features();
            // active parts of the structures that have active atoms
            vector3d ax_a, pos_a;
            vector2i ty_a;
            std::vector<int> keep;
            std::vector<vector1i> atoms;
            for (int s = 0; s < n_st; ++s) {
                vector1i sel, t;
                for (int i = 0; i < (int)types[s].size(); ++i) {
                    if (full) {
                        sel.push_back(i), t.push_back(types[s][i]);
                        continue;
                    }
                    auto it = std::find(idx.begin(), idx.end(), types[s][i]);
                    if (it != idx.end()) sel.push_back(i), t.push_back((int)(it - idx.begin()));
                }
                if (sel.empty()) continue;
                vector2d p(3, vector1d(sel.size()));
                for (int a = 0; a < 3; ++a)
                    for (size_t j = 0; j < sel.size(); ++j) p[a][j] = positions_c[s][a][sel[j]];
                keep.push_back(s), atoms.push_back(sel);
                ax_a.push_back(axis[s]), pos_a.push_back(p), ty_a.push_back(t);
            }
            if (keep.empty()) continue;
            Packed pk = pack(ax_a, pos_a, ty_a);
            const int m = (int)keep.size();
            plm_batch* b = nullptr;
            check(plm_batch_create(models[n]->m, m, pk.axes.data(), pk.pos.data(), pk.types.data(),
                                   pk.n_atoms.data(), &b));
            long rows = 0;
            plm_batch_n_rows(b, &rows);
            std::vector<double> xs((size_t)rows * Fn);
            const int rc = plm_features_x(models[n]->m, b, PLM_FP64, xs.data());
            plm_batch_destroy(b);
            check(rc);
            auto copy_row = [&](long dst, long src) {
                for (int c = 0; c < Fn; ++c) X(dst, c0 + c) = xs[(size_t)src * Fn + c];
            };
            long f_src = 7L * m;
            for (int q = 0; q < m; ++q) {
                const int s = keep[q];
                copy_row(s, q);
                if (force_st[s]) {
                    for (int k = 0; k < 6; ++k) copy_row(srow[s] + k, m + 6L * q + k);
                    for (size_t j = 0; j < atoms[q].size(); ++j)
                        for (int a = 0; a < 3; ++a)
                            copy_row(frow[s] + 3L * atoms[q][j] + a, f_src + 3L * j + a);
                }
                f_src += 3L * (long)atoms[q].size();
            }
        }
    }
    py::array_t<double, py::array::f_style>& get_x() { return x_; }
    const vector1i& get_fbegin() const { return xf_begin_; }
    const vector1i& get_sbegin() const { return xs_begin_; }
    const vector1i& get_cumulative_n_features() const { return cum_; }
    const vector1i& get_n_data() const { return n_data_; }

  private:
    py::array_t<double, py::array::f_style> x_;
    vector1i xf_begin_, xs_begin_, n_data_, cum_;
};

// FeaturesAttr (py_features_attr.cpp:11-66): column attributes of a model
class FeaturesAttr {
  public:
    explicit FeaturesAttr(const py::dict& params) {
        ModelOwner mo(params);
        int nl = 0, np = 0;
        check(plm_model_feature_attrs(mo.m, &nl, &np, nullptr, nullptr, nullptr, nullptr, nullptr));
        const int nt = params["n_type"].cast<int>();
        const bool gtinv = params["model"].cast<py::dict>()["feature_type"].cast<std::string>() == "gtinv";
        vector1i radial(nl), gt(nl), tc(6 * (size_t)nl), poly(3 * (size_t)np), tps((size_t)nt * nt);
        check(plm_model_feature_attrs(mo.m, nullptr, nullptr, radial.data(), gt.data(), tc.data(),
                                      poly.data(), tps.data()));
        radial_ = radial;
        if (gtinv) gtinv_ = gt;
        for (int f = 0; f < nl; ++f) {
            vector1i c;
            for (int k = 0; k < 6; ++k)
                if (tc[6 * f + k] >= 0) c.push_back(tc[6 * f + k]);
            tcomb_.push_back(c);
        }
        for (int p = 0; p < np; ++p) {
            vector1i c;
            for (int k = 0; k < 3; ++k)
                if (poly[3 * p + k] >= 0) c.push_back(poly[3 * p + k]);
            poly_.push_back(c);
        }
        type_pairs_.assign(nt, vector1i(nt));
        for (int a = 0; a < nt; ++a)
            for (int b = 0; b < nt; ++b) type_pairs_[a][b] = tps[a * nt + b];
    }
    const vector1i& get_radial_ids() const { return radial_; }
    const vector1i& get_gtinv_ids() const { return gtinv_; }
    const vector2i& get_tcomb_ids() const { return tcomb_; }
    const vector2i& get_polynomial_ids() const { return poly_; }
    const vector2i& get_type_pairs() const { return type_pairs_; }

  private:
    vector1i radial_, gtinv_;
    vector2i tcomb_, poly_, type_pairs_;
};

// Readgtinv (polymlp_read_gtinv.cpp:23-59): the coupling tables kept for
// (order, maxl, version), read from the same version-1 tables the params
// dicts are built from (pypolymlp_amd.gtinv)
class Readgtinv {
  public:
    Readgtinv(int order, const vector1i& maxl, int version) {
        py::object r = py::module_::import("pypolymlp_amd.gtinv")
                           .attr("get_gtinv_attrs")(order, maxl, version);
        py::tuple t = r.cast<py::tuple>();
        l_comb_ = t[0].cast<vector2i>();
        lm_seq_ = t[1].cast<std::vector<vector2i>>();
        lm_coeffs_ = t[2].cast<vector2d>();
    }
    const std::vector<vector2i>& get_lm_seq() const { return lm_seq_; }
    const vector2i& get_l_comb() const { return l_comb_; }
    const vector2d& get_lm_coeffs() const { return lm_coeffs_; }

  private:
    std::vector<vector2i> lm_seq_;
    vector2i l_comb_;
    vector2d lm_coeffs_;
};

// PotentialPropertiesFast (py_properties_fast.cpp:10-85)
class PotentialPropertiesFast {
  public:
    PotentialPropertiesFast(const py::dict& params, const vector1d& coeffs) : mo_(params) {
        check(plm_model_set_coeffs(mo_.m, coeffs.data(), (int)coeffs.size()));
    }
    void eval(const vector2d& axis, const vector2d& positions_c, const vector1i& types,
              const bool /*use_openmp*/) {
        Packed p = pack({axis}, {positions_c}, {types});
        const int n = p.n_atoms[0];
        std::vector<double> f((size_t)3 * n), s(6);
        check(plm_eval(mo_.m, p.axes.data(), p.pos.data(), p.types.data(), n, PLM_FP64, &energy_,
                       f.data(), s.data()));
        force_.assign(n, vector1d(3));
        for (int i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) force_[i][a] = f[3 * i + a];
        stress_ = s;
    }
    void eval_multiple(const vector3d& axis, const vector3d& positions_c, const vector2i& types) {
        Packed p = pack(axis, positions_c, types);
        const int n_st = (int)p.n_atoms.size();
        plm_batch* b = nullptr;
        check(plm_batch_create(mo_.m, n_st, p.axes.data(), p.pos.data(), p.types.data(),
                               p.n_atoms.data(), &b));
        long A = 0;
        for (int n : p.n_atoms) A += n;
        std::vector<double> e(n_st), f((size_t)3 * A), s((size_t)6 * n_st);
        const int rc = plm_eval_batch(mo_.m, b, PLM_FP64, e.data(), f.data(), s.data());
        plm_batch_destroy(b);
        check(rc);
        e_array_ = e;
        f_array_.assign(n_st, {});
        s_array_.assign(n_st, vector1d(6));
        long off = 0;
        for (int k = 0; k < n_st; ++k) {
            f_array_[k].assign(p.n_atoms[k], vector1d(3));
            for (int i = 0; i < p.n_atoms[k]; ++i)
                for (int a = 0; a < 3; ++a) f_array_[k][i][a] = f[3 * (off + i) + a];
            off += p.n_atoms[k];
            for (int a = 0; a < 6; ++a) s_array_[k][a] = s[6 * k + a];
        }
    }
    const double& get_e() const { return energy_; }
    const vector2d& get_f() const { return force_; }
    const vector1d& get_s() const { return stress_; }
    const vector1d& get_e_array() const { return e_array_; }
    const vector3d& get_f_array() const { return f_array_; }
    const vector2d& get_s_array() const { return s_array_; }

  private:
    ModelOwner mo_;
    double energy_ = 0.0;
    vector2d force_;
    vector1d stress_;
    vector1d e_array_;
    vector3d f_array_;
    vector2d s_array_;
};

// NeighborFull / NeighborHalf / NeighborCell accessors used by the reference tests
// (neighbor_full.cpp:79-126, neighbor_half.cpp:103-124, neighbor_cell.h:78-82).
struct NeighborList {
    vector1i offset, neigh;
    vector1d d;
    NeighborList(const vector2d& axis, const vector2d& positions_c, double cutoff, bool half) {
        Packed p = pack({axis}, {positions_c}, {vector1i(positions_c[0].size(), 0)});
        const int n = p.n_atoms[0];
        long m = 0;
        check(plm_neighbor(p.axes.data(), p.pos.data(), n, cutoff, half ? 1 : 0, nullptr, nullptr,
                           nullptr, nullptr, 0, &m));
        offset.resize(n + 1), neigh.resize(m), d.resize(3 * m);
        vector1i tid(m);
        check(plm_neighbor(p.axes.data(), p.pos.data(), n, cutoff, half ? 1 : 0, offset.data(),
                           neigh.data(), tid.data(), d.data(), m, &m));
    }
};

class NeighborFullPy {
  public:
    NeighborFullPy(const vector2d& axis, const vector2d& positions_c, double cutoff)
        : nl_(axis, positions_c, cutoff, false) {}
    std::vector<vector2d> get_dis_array(int n_type, const vector1i& types) {
        const int n = (int)nl_.offset.size() - 1;
        std::vector<vector2d> out(n, vector2d(n_type));
        for (int i = 0; i < n; ++i)
            for (int k = nl_.offset[i]; k < nl_.offset[i + 1]; ++k) {
                const double x = nl_.d[3 * k], y = nl_.d[3 * k + 1], z = nl_.d[3 * k + 2];
                out[i][types[nl_.neigh[k]]].push_back(std::sqrt(x * x + y * y + z * z));
            }
        return out;
    }
    std::vector<std::vector<vector2d>> get_diff_array(int n_type, const vector1i& types) {
        const int n = (int)nl_.offset.size() - 1;
        std::vector<std::vector<vector2d>> out(n, std::vector<vector2d>(n_type));
        for (int i = 0; i < n; ++i)
            for (int k = nl_.offset[i]; k < nl_.offset[i + 1]; ++k)
                out[i][types[nl_.neigh[k]]].push_back({nl_.d[3 * k], nl_.d[3 * k + 1], nl_.d[3 * k + 2]});
        return out;
    }
    std::vector<vector2i> get_atom2_array(int n_type, const vector1i& types) {
        const int n = (int)nl_.offset.size() - 1;
        std::vector<vector2i> out(n, vector2i(n_type));
        for (int i = 0; i < n; ++i)
            for (int k = nl_.offset[i]; k < nl_.offset[i + 1]; ++k)
                out[i][types[nl_.neigh[k]]].push_back(nl_.neigh[k]);
        return out;
    }

  private:
    NeighborList nl_;
};

class NeighborHalfPy {
  public:
    NeighborHalfPy(const vector2d& axis, const vector2d& positions_c, double cutoff, bool)
        : nl_(axis, positions_c, cutoff, true) {}
    vector3d get_diff_list() {
        const int n = (int)nl_.offset.size() - 1;
        vector3d out(n);
        for (int i = 0; i < n; ++i)
            for (int k = nl_.offset[i]; k < nl_.offset[i + 1]; ++k)
                out[i].push_back({nl_.d[3 * k], nl_.d[3 * k + 1], nl_.d[3 * k + 2]});
        return out;
    }
    vector2i get_half_list() {
        const int n = (int)nl_.offset.size() - 1;
        vector2i out(n);
        for (int i = 0; i < n; ++i)
            for (int k = nl_.offset[i]; k < nl_.offset[i + 1]; ++k) out[i].push_back(nl_.neigh[k]);
        return out;
    }

  private:
    NeighborList nl_;
};

class NeighborCellPy {
  public:
    NeighborCellPy(const vector2d& axis, const vector2d& positions_c, double cutoff) {
        Packed p = pack({axis}, {positions_c}, {vector1i(positions_c[0].size(), 0)});
        const int n = p.n_atoms[0];
        int nt = 0;
        check(plm_cell_translations(p.axes.data(), p.pos.data(), n, cutoff, nullptr, 0, nullptr, &nt));
        std::vector<double> tr((size_t)3 * nt), pr((size_t)3 * n);
        check(plm_cell_translations(p.axes.data(), p.pos.data(), n, cutoff, tr.data(), nt, pr.data(), &nt));
        for (int k = 0; k < nt; ++k) trans_.push_back({tr[3 * k], tr[3 * k + 1], tr[3 * k + 2]});
        pos_.assign(3, vector1d(n));
        for (int a = 0; a < 3; ++a)
            for (int j = 0; j < n; ++j) pos_[a][j] = pr[(size_t)a * n + j];
    }
    const vector2d& get_translations() const { return trans_; }
    const vector2d& get_positions_cartesian() const { return pos_; }

  private:
    vector2d trans_, pos_;
};

}  // namespace

PYBIND11_MODULE(libmlpcpp, m) {
    m.doc() = "MI355X hot path of pypolymlp (C-ABI: include/polymlp_amd.h)";
    py::class_<PotentialModel>(m, "PotentialModel")
        .def(py::init<const py::dict&, const vector3d&, const vector3d&, const vector2i&,
                      const vector1i&, const std::vector<bool>&, const vector1i&>())
        .def("get_x", &PotentialModel::get_x, py::return_value_policy::reference_internal)
        .def("get_fbegin", &PotentialModel::get_fbegin, py::return_value_policy::reference_internal)
        .def("get_sbegin", &PotentialModel::get_sbegin, py::return_value_policy::reference_internal)
        .def("get_n_data", &PotentialModel::get_n_data, py::return_value_policy::reference_internal);
    py::class_<PotentialHybridModel>(m, "PotentialHybridModel")
        .def(py::init<const std::vector<py::dict>&, const vector3d&, const vector3d&,
                      const vector2i&, const vector1i&, const std::vector<bool>&,
                      const vector1i&>())
        .def("get_x", &PotentialHybridModel::get_x, py::return_value_policy::reference_internal)
        .def("get_fbegin", &PotentialHybridModel::get_fbegin,
             py::return_value_policy::reference_internal)
        .def("get_sbegin", &PotentialHybridModel::get_sbegin,
             py::return_value_policy::reference_internal)
        .def("get_cumulative_n_features", &PotentialHybridModel::get_cumulative_n_features,
             py::return_value_policy::reference_internal)
        .def("get_n_data", &PotentialHybridModel::get_n_data,
             py::return_value_policy::reference_internal);
    py::class_<FeaturesAttr>(m, "FeaturesAttr")
        .def(py::init<const py::dict&>())
        .def("get_radial_ids", &FeaturesAttr::get_radial_ids, py::return_value_policy::reference_internal)
        .def("get_gtinv_ids", &FeaturesAttr::get_gtinv_ids, py::return_value_policy::reference_internal)
        .def("get_tcomb_ids", &FeaturesAttr::get_tcomb_ids, py::return_value_policy::reference_internal)
        .def("get_polynomial_ids", &FeaturesAttr::get_polynomial_ids,
             py::return_value_policy::reference_internal)
        .def("get_type_pairs", &FeaturesAttr::get_type_pairs, py::return_value_policy::reference_internal);
    py::class_<Readgtinv>(m, "Readgtinv")
        .def(py::init<int, const vector1i&, int>())
        .def("get_lm_seq", &Readgtinv::get_lm_seq, py::return_value_policy::reference_internal)
        .def("get_l_comb", &Readgtinv::get_l_comb, py::return_value_policy::reference_internal)
        .def("get_lm_coeffs", &Readgtinv::get_lm_coeffs, py::return_value_policy::reference_internal);
    py::class_<PotentialPropertiesFast>(m, "PotentialPropertiesFast")
        .def(py::init<const py::dict&, const vector1d&>())
        .def("eval", &PotentialPropertiesFast::eval)
        .def("eval_multiple", &PotentialPropertiesFast::eval_multiple)
        .def("get_e", &PotentialPropertiesFast::get_e)
        .def("get_f", &PotentialPropertiesFast::get_f)
        .def("get_s", &PotentialPropertiesFast::get_s)
        .def("get_e_array", &PotentialPropertiesFast::get_e_array)
        .def("get_f_array", &PotentialPropertiesFast::get_f_array)
        .def("get_s_array", &PotentialPropertiesFast::get_s_array);
    py::class_<NeighborFullPy>(m, "NeighborFull")
        .def(py::init<const vector2d&, const vector2d&, double>())
        .def("get_distances", &NeighborFullPy::get_dis_array)
        .def("get_differences", &NeighborFullPy::get_diff_array)
        .def("get_neighbor_indices", &NeighborFullPy::get_atom2_array);
    py::class_<NeighborHalfPy>(m, "NeighborHalf")
        .def(py::init<const vector2d&, const vector2d&, double, bool>())
        .def("get_differences", &NeighborHalfPy::get_diff_list)
        .def("get_neighbor_indices", &NeighborHalfPy::get_half_list);
    py::class_<NeighborCellPy>(m, "NeighborCell")
        .def(py::init<const vector2d&, const vector2d&, double>())
        .def("get_translations", &NeighborCellPy::get_translations)
        .def("get_positions_cartesian", &NeighborCellPy::get_positions_cartesian);
}
