// CPU oracle harness.
//
// This file is test infrastructure. It links the reference's own C++
// sources for the hot path (compiled in place by oracle/Makefile) and adds a
// restatement of the two pieces that cannot be compiled here because they
// need Eigen: the polynomial X-row assembly of Model (model.cpp) and the row
// layout of PyModel (py_model.cpp). Only tests, smoke() and the CPU-baseline
// leg of bench.py call into it, and only as the checker.
//
// Reference pieces used as-is:
//   NeighborCell / NeighborFull / NeighborHalf  src/cxx/src/compute/neighbor_*.cpp
//   Local::gtinv_d, LocalPair::pair_d            src/cxx/src/compute/local*.cpp
//   PolymlpAPI / Features / Potential            src/cxx/src/polymlp/*.cpp
//   PolymlpEval                                  src/cxx/src/compute/polymlp_eval.cpp
//   parse_polymlp (polymlp.yaml reader)          src/cxx/src/polymlp/polymlp_parse_potential.cpp
//   get_fn_ / get_ylm_                           src/cxx/src/polymlp/polymlp_functions_interface.cpp
//
// Restated here:
//   Model::model_polynomial / model_order{1,2,3}  model.cpp:159-267
//   Model::reshape                                model.cpp:119-157
//   PyModel::PyModel / set_index row layout       py_model.cpp:10-106

#include <cstring>
#include <memory>
#include <stdexcept>

#include "compute/local.h"
#include "compute/local_pair.h"
#include "compute/neighbor_cell.h"
#include "compute/neighbor_full.h"
#include "compute/neighbor_half.h"
#include "compute/polymlp_eval.h"
#include "polymlp/polymlp_api.h"
#include "polymlp/polymlp_functions_interface.h"
#include "polymlp/polymlp_parse_potential.h"
#include "polymlp/polymlp_parse_potential_legacy.h"
#include "polymlp/polymlp_read_gtinv.h"

namespace {

struct RefModel {
    feature_params fp;
    vector1d coeffs;  // as stored in polymlp.yaml (coeffs / scales)
    vector1d pot2;    // 2 * coeffs / scales, as parse_polymlp returns them
    PolymlpAPI feat;  // features with derivatives (PolymlpAPI::set_features)
    std::unique_ptr<PolymlpEval> eval;
    std::vector<std::vector<PolynomialTerm>> terms1, terms2, terms3;
    int n_features = 0;
};

vector2d axis_from(const double* a) {
    // a is row-major 3x3 with lattice vectors as columns (axis[i][col]).
    vector2d axis(3, vector1d(3));
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) axis[i][j] = a[3 * i + j];
    return axis;
}

vector2d pos_from(const double* p, int n) {
    // p is row-major 3 x n (Cartesian components as rows).
    vector2d pos(3, vector1d(n));
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < n; ++j) pos[i][j] = p[(size_t)i * n + j];
    return pos;
}

// model.cpp:181-267 restated on plain row-major buffers.
// de[L]; df[(3N) x L] row-major (row 3k+a); ds[6 x L]
void model_polynomial(const RefModel& m, int type1, bool force, int n_atom,
                      const vector1d& de, const std::vector<double>& df,
                      const std::vector<double>& ds, int L, double* xe,
                      double* xf, double* xs, int F) {
    const int n3 = 3 * n_atom;
    for (const auto& term : m.terms1[type1]) {
        const int col = term.global_id, c1 = term.local_ids[0];
        xe[col] += de[c1];
        if (!force) continue;
        for (int k = 0; k < n3; ++k) xf[(size_t)k * F + col] += df[(size_t)k * L + c1];
        for (int k = 0; k < 6; ++k) xs[(size_t)k * F + col] += ds[(size_t)k * L + c1];
    }
    for (const auto& term : m.terms2[type1]) {
        const int col = term.global_id;
        const int c1 = term.local_ids[0], c2 = term.local_ids[1];
        xe[col] += de[c1] * de[c2];
        if (!force) continue;
        const double v1 = de[c2], v2 = de[c1];
        for (int k = 0; k < n3; ++k)
            xf[(size_t)k * F + col] += v1 * df[(size_t)k * L + c1] + v2 * df[(size_t)k * L + c2];
        for (int k = 0; k < 6; ++k)
            xs[(size_t)k * F + col] += v1 * ds[(size_t)k * L + c1] + v2 * ds[(size_t)k * L + c2];
    }
    for (const auto& term : m.terms3[type1]) {
        const int col = term.global_id;
        const int c1 = term.local_ids[0], c2 = term.local_ids[1], c3 = term.local_ids[2];
        xe[col] += de[c1] * de[c2] * de[c3];
        if (!force) continue;
        const double v1 = de[c2] * de[c3], v2 = de[c1] * de[c3], v3 = de[c1] * de[c2];
        for (int k = 0; k < n3; ++k)
            xf[(size_t)k * F + col] += v1 * df[(size_t)k * L + c1] + v2 * df[(size_t)k * L + c2] +
                                      v3 * df[(size_t)k * L + c3];
        for (int k = 0; k < 6; ++k)
            xs[(size_t)k * F + col] += v1 * ds[(size_t)k * L + c1] + v2 * ds[(size_t)k * L + c2] +
                                      v3 * ds[(size_t)k * L + c3];
    }
}

// Model::run + Model::gtinv/pair + reshape (model.cpp:39-157), writing the
// structure's energy row xe[F], stress rows xs[6xF] and force rows xf[3N x F].
void structure_rows(RefModel& m, const vector2d& axis, const vector2d& pos,
                    const vector1i& types, bool force, double* xe, double* xf,
                    double* xs) {
    const int F = m.n_features;
    const int n_atom = (int)types.size();
    NeighborFull neigh(axis, pos, m.fp.cutoff);
    std::fill(xe, xe + F, 0.0);
    if (force) {
        std::fill(xf, xf + (size_t)3 * n_atom * F, 0.0);
        std::fill(xs, xs + (size_t)6 * F, 0.0);
    }
    const bool is_pair = (m.fp.feature_type == "pair");
    Local local(n_atom);
    LocalPair local_pair(n_atom);
    std::vector<double> df, ds;
    for (int atom1 = 0; atom1 < n_atom; ++atom1) {
        const int type1 = types[atom1];
        vector1d de;
        vector2d dfx, dfy, dfz, dsv;
        if (is_pair) {
            if (force) local_pair.pair_d(m.feat, neigh, types, atom1, de, dfx, dfy, dfz, dsv);
            else local_pair.pair(m.feat, neigh, types, atom1, de);
        } else {
            if (force) local.gtinv_d(m.feat, neigh, types, atom1, de, dfx, dfy, dfz, dsv);
            else local.gtinv(m.feat, neigh, types, atom1, de);
        }
        const int L = (int)de.size();
        if (force) {
            df.assign((size_t)3 * n_atom * L, 0.0);
            ds.assign((size_t)6 * L, 0.0);
            for (int i = 0; i < L; ++i) {
                for (int j = 0; j < n_atom; ++j) {
                    df[(size_t)(3 * j) * L + i] = dfx[i][j];
                    df[(size_t)(3 * j + 1) * L + i] = dfy[i][j];
                    df[(size_t)(3 * j + 2) * L + i] = dfz[i][j];
                }
                for (int j = 0; j < 6; ++j) ds[(size_t)j * L + i] = dsv[i][j];
            }
        }
        model_polynomial(m, type1, force, n_atom, de, df, ds, L, xe, xf, xs, F);
    }
}

thread_local std::string g_err;

}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

// Load a polymlp.yaml (gtinv tables come from the .bin files beside this
// library, as in the reference). with_eval builds PolymlpEval as well.
void* ref_model_load(const char* yaml, int with_eval) {
    try {
        auto m = std::make_unique<RefModel>();
        vector1d pot2;
        std::vector<std::string> ele;
        vector1d mass;
        // both formats give pot2 = 2 * coeffs / scales (polymlp_api.cpp:20-32)
        if (check_polymlp_legacy(yaml))
            parse_polymlp_legacy(yaml, m->fp, pot2, ele, mass);
        else
            parse_polymlp(yaml, m->fp, pot2, ele, mass);
        m->pot2 = pot2;
        m->coeffs.resize(pot2.size());
        for (size_t i = 0; i < pot2.size(); ++i) m->coeffs[i] = 0.5 * pot2[i];
        m->feat.set_features(m->fp);
        m->n_features = m->feat.get_n_variables();
        auto& maps = m->feat.get_maps();
        m->terms1.resize(m->fp.n_type);
        m->terms2.resize(m->fp.n_type);
        m->terms3.resize(m->fp.n_type);
        for (int t = 0; t < m->fp.n_type; ++t) {
            for (const auto& term : maps.maps_type[t].polynomial) {
                if (term.local_ids.size() == 1) m->terms1[t].push_back(term);
                else if (term.local_ids.size() == 2) m->terms2[t].push_back(term);
                else if (term.local_ids.size() == 3) m->terms3[t].push_back(term);
            }
        }
        if (with_eval) {
            if ((int)m->coeffs.size() != m->n_features)
                throw std::runtime_error("n_coeffs does not match the number of features");
            m->eval = std::make_unique<PolymlpEval>(m->fp, m->coeffs);
        }
        return m.release();
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void ref_model_free(void* h) { delete static_cast<RefModel*>(h); }

int ref_model_n_features(void* h) { return static_cast<RefModel*>(h)->n_features; }

// Number of linear features for atom type t (size of the per-atom feature
// vector dn of Local::gtinv).
int ref_model_n_linear(void* h, int type1) {
    auto* m = static_cast<RefModel*>(h);
    return (int)m->terms1[type1].size();
}

// Translations and (possibly rewrapped) positions of NeighborCell.
// Returns the number of translations; fills trans[n_T*3] if non-null and
// pos_rev[3*n] if non-null.
long ref_neighbor_cell(const double* axis9, const double* pos, int n, double cutoff,
                       double* trans, long max_trans, double* pos_rev, double* axis_out) {
    NeighborCell cell(axis_from(axis9), pos_from(pos, n), cutoff);
    const auto& tr = cell.get_translations();
    if (trans) {
        for (long k = 0; k < (long)tr.size() && k < max_trans; ++k)
            for (int a = 0; a < 3; ++a) trans[3 * k + a] = tr[k][a];
    }
    if (pos_rev) {
        const auto& p = cell.get_positions_cartesian();
        for (int a = 0; a < 3; ++a)
            for (int j = 0; j < n; ++j) pos_rev[(size_t)a * n + j] = p[a][j];
    }
    if (axis_out) {
        const auto& ax = cell.get_axis();
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) axis_out[3 * a + b] = ax[a][b];
    }
    return (long)tr.size();
}

// Full / half neighbour lists. Returns the number of pairs; when the output
// pointers are non-null they receive offset[n+1], neigh[M], d[M*3] with
// d = r_j - r_i + T (full list) or the half-list convention of
// neighbor_half.cpp (also r_j - r_i + T; PolymlpEval negates it).
long ref_neighbor(const double* axis9, const double* pos, int n, double cutoff, int half,
                  int* offset, int* neigh, double* d) {
    auto axis = axis_from(axis9);
    auto p = pos_from(pos, n);
    auto fill = [&](auto& nb) -> long {
        long total = 0;
        for (int i = 0; i < n; ++i) total += nb.size(i);
        if (offset) {
            long c = 0;
            offset[0] = 0;
            for (int i = 0; i < n; ++i) {
                auto [b, e] = nb.range(i);
                for (int k = b; k < e; ++k) {
                    if (neigh) neigh[c] = nb.neighbor_atom(k);
                    if (d) {
                        double x, y, z;
                        nb.diff_ji(k, x, y, z);
                        d[3 * c] = x, d[3 * c + 1] = y, d[3 * c + 2] = z;
                    }
                    ++c;
                }
                offset[i + 1] = (int)c;
            }
        }
        return total;
    };
    if (half) {
        // OpenMP over atoms: the per-atom lists are built independently and
        // concatenated in atom order (neighbor_half.cpp:30-94), as serially
        NeighborHalf nb(axis, p, cutoff, true);
        return fill(nb);
    }
    NeighborFull nb(axis, p, cutoff);
    return fill(nb);
}

// X rows of one structure: xe[F], xs[6*F], xf[3n*F] (row-major, force row 3k+a).
int ref_structure_rows(void* h, const double* axis9, const double* pos, const int* types,
                       int n, int force, double* xe, double* xs, double* xf) {
    try {
        auto* m = static_cast<RefModel*>(h);
        vector1i t(types, types + n);
        structure_rows(*m, axis_from(axis9), pos_from(pos, n), t, force != 0, xe, xf, xs);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// PyModel over a batch of structures (py_model.cpp:10-54), all with forces:
// X (total_rows x F) row-major with the reference's row layout
// [n_st energy rows | 6 stress rows per structure | 3N force rows per structure].
// Structures are processed in parallel (OpenMP over structures, py_model.cpp:41).
int ref_x_batch(void* h, int n_st, const double* axes, const double* pos_concat,
                const int* types_concat, const int* n_atoms, int force, double* X) {
    auto* m = static_cast<RefModel*>(h);
    const int F = m->n_features;
    std::vector<long> atom_off(n_st + 1, 0), frow(n_st, 0), srow(n_st, 0);
    for (int s = 0; s < n_st; ++s) atom_off[s + 1] = atom_off[s] + n_atoms[s];
    long r = n_st;
    if (force) {
        for (int s = 0; s < n_st; ++s) srow[s] = r, r += 6;
        for (int s = 0; s < n_st; ++s) frow[s] = r, r += 3L * n_atoms[s];
    }
    int err = 0;
#pragma omp parallel for schedule(guided, 1)
    for (int s = 0; s < n_st; ++s) {
        const int n = n_atoms[s];
        const double* p = pos_concat + 3 * atom_off[s];
        vector1i t(types_concat + atom_off[s], types_concat + atom_off[s] + n);
        try {
            structure_rows(*m, axis_from(axes + 9 * s), pos_from(p, n), t, force != 0,
                           X + (size_t)s * F, force ? X + (size_t)frow[s] * F : nullptr,
                           force ? X + (size_t)srow[s] * F : nullptr);
        } catch (...) {
#pragma omp atomic write
            err = 1;
        }
    }
    return err;
}

// PolymlpEval::eval through NeighborHalf (py_properties_fast.cpp:19-28).
int ref_eval(void* h, const double* axis9, const double* pos, const int* types, int n,
             int use_openmp, double* energy, double* forces, double* stress) {
    try {
        auto* m = static_cast<RefModel*>(h);
        if (!m->eval) throw std::runtime_error("model loaded without eval");
        vector1i t(types, types + n);
        NeighborHalf neigh(axis_from(axis9), pos_from(pos, n), m->fp.cutoff, use_openmp != 0);
        double e;
        vector2d f;
        vector1d s;
        m->eval->eval(t, neigh, use_openmp != 0, e, f, s);
        *energy = e;
        for (int i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) forces[3 * i + a] = f[i][a];
        for (int k = 0; k < 6; ++k) stress[k] = s[k];
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// get_fn_ (polymlp_functions_interface.cpp:41-63) at n distances: fn, fn_d
// (n x n_params, row-major).
void ref_get_fn(int n, const double* r, const double* params, int n_params, double cutoff,
                double* fn, double* fn_d) {
    feature_params fp;
    fp.pair_type = "gaussian";
    fp.cutoff = cutoff;
    vector2d par(n_params, vector1d(2));
    for (int k = 0; k < n_params; ++k) par[k] = {params[2 * k], params[2 * k + 1]};
    for (int i = 0; i < n; ++i) {
        vector1d f, d;
        get_fn_(r[i], fp, par, f, d);
        for (int k = 0; k < n_params; ++k) {
            fn[(size_t)i * n_params + k] = f[k];
            fn_d[(size_t)i * n_params + k] = d[k];
        }
    }
}

// get_ylm_ (polymlp_functions_interface.cpp:113-139) at n points with
// r = |xyz|: ylm (n x nh complex pairs), dylm (n x nh x 3 complex pairs).
void ref_get_ylm(int n, const double* xyz, int lmax, double* ylm, double* dylm) {
    const int nh = (lmax + 1) * (lmax + 2) / 2;
    for (int i = 0; i < n; ++i) {
        const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        const double r = std::sqrt(x * x + y * y + z * z);
        vector1dc Y, dx, dy, dz;
        get_ylm_(r, x, y, z, lmax, Y, dx, dy, dz);
        for (int k = 0; k < nh; ++k) {
            const size_t o = (size_t)i * nh + k;
            ylm[2 * o] = Y[k].real(), ylm[2 * o + 1] = Y[k].imag();
            const dc* d3[3] = {&dx[k], &dy[k], &dz[k]};
            for (int a = 0; a < 3; ++a) {
                dylm[2 * (3 * o + a)] = d3[a]->real();
                dylm[2 * (3 * o + a) + 1] = d3[a]->imag();
            }
        }
    }
}

// PolymlpAPI::compute_anlmtp_conjugate + compute_features + (with coefficients)
// compute_sum_of_prod_anlmtp (polymlp_api.cpp:63-185) for one atom of type1
// from its non-conjugate order parameters. Sizes are returned through n_out:
// [n_nlm, n_features, n_heads]; outputs may be null.
int ref_api(void* h, int type1, const double* re, const double* im, int n_noconj, double* anlm,
            double* features, double* prod_e, double* prod_f, int* n_out) {
    try {
        auto* m = static_cast<RefModel*>(h);
        vector1d r(re, re + n_noconj), i(im, im + n_noconj);
        vector1dc a;
        m->feat.compute_anlmtp_conjugate(r, i, type1, a);
        vector1d f;
        m->feat.compute_features(a, type1, f);
        n_out[0] = (int)a.size();
        n_out[1] = (int)f.size();
        n_out[2] = 0;
        if (anlm)
            for (size_t k = 0; k < a.size(); ++k) anlm[2 * k] = a[k].real(), anlm[2 * k + 1] = a[k].imag();
        if (features) std::copy(f.begin(), f.end(), features);
        if (prod_e || prod_f) {
            PolymlpAPI api;
            api.set_potential_model(m->fp, m->pot2);
            vector1dc pe, pf;
            api.compute_sum_of_prod_anlmtp(a, type1, pe, pf);
            n_out[2] = (int)pe.size();
            for (size_t k = 0; k < pe.size(); ++k) {
                if (prod_e) prod_e[2 * k] = pe[k].real(), prod_e[2 * k + 1] = pe[k].imag();
                if (prod_f) prod_f[2 * k] = pf[k].real(), prod_f[2 * k + 1] = pf[k].imag();
            }
        }
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// Readgtinv (polymlp_read_gtinv.cpp:12-59): flattened l_comb / lm_seq / coeffs.
// Returns the number of l-combinations. n_terms[i] = number of m-terms of
// invariant i; lm flat (sum(n_terms) * order_i); coeffs flat.
long ref_gtinv(int order, const int* maxl, int n_maxl, int version, int* orders,
               int* l_flat, int* n_terms, int* lm_flat, double* coeffs) {
    vector1i ml(maxl, maxl + n_maxl);
    Readgtinv rg(order, ml, version);
    const auto& lc = rg.get_l_comb();
    const auto& lm = rg.get_lm_seq();
    const auto& cf = rg.get_lm_coeffs();
    long pl = 0, pm = 0, pc = 0;
    for (size_t i = 0; i < lc.size(); ++i) {
        if (orders) orders[i] = (int)lc[i].size();
        if (n_terms) n_terms[i] = (int)lm[i].size();
        for (int v : lc[i]) { if (l_flat) l_flat[pl] = v; ++pl; }
        for (const auto& row : lm[i])
            for (int v : row) { if (lm_flat) lm_flat[pm] = v; ++pm; }
        for (double v : cf[i]) { if (coeffs) coeffs[pc] = v; ++pc; }
    }
    return (long)lc.size();
}

}  // extern "C"
