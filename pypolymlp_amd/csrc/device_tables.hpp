// Plain-old-data views of the model and structure-batch tables in device
// memory. Passed by value to kernels. All index arrays are int32.
#pragma once

#include <cstdint>

namespace plm {

constexpr int kMaxTypes = 8;
constexpr int kMaxFn = 64;      // radial functions per type pair
constexpr int kMaxLmHalf = 120; // (lmax+1)(lmax+2)/2 for lmax <= 14

// Per-type tables (offsets into the flat arrays of ModelDev).
struct TypeDev {
    int n_nlm;        // local order-parameter components (incl. conjugates)
    int n_noconj;     // non-conjugate components
    int n_lin;        // linear features valid for this type
    // a_full fill: local component k = conj ? cc * conj(a_nc[src]) : a_nc[src]
    const int* nlm_src;        // [n_nlm] noconj index of the source
    const float* nlm_sign;     // [n_nlm] +1 plain, cc for conjugates (sign of imag flips)
    const unsigned char* nlm_conj;  // [n_nlm]
    // anlm accumulation grouped by type pair: heads with tp == p
    const int* nc_tp_off;      // [n_tp+1]
    const int* nc_tp_head;     // noconj index
    const int* nc_tp_nid;      // radial n_id
    const int* nc_tp_ylm;      // ylm key
    // feature values: x_f = sum coeff * Re(prod a[keys])
    const int* fv_off;         // [n_lin+1]
    const double* fv_coeff;
    const int* fv_keys;        // [nterm*4], -1 padded
    int fv_skip;               // 1: terms with |Re prod| <= 1e-20 are dropped (more than 1000
                               // value products, polymlp_features.cpp:184-193)
    // derivative weights W'[f, h]: nonzeros sorted by (f, tp)
    const int* wz_off;         // [n_lin*n_tp + 1]
    const int* wz_nid;         // radial n_id of the head
    const int* wz_ylm;         // ylm key of the head
    const int* wz_term_off;    // [nnz+1]
    const double* wt_coeff;    // folded coefficient (cc applied for conjugate heads)
    const int* wt_keys;        // [nterm*3], -1 padded: product of all but the head
    const unsigned char* wt_conj;  // product enters conjugated
    // The same terms laid out for the W' / value kernel: items (W' nonzeros,
    // features) in order of decreasing term count, in blocks of 64 (one per
    // lane); term j of the block's lane l at row j, so a wave's loads of one
    // row are contiguous. Rows past an item's count are padding (keys -1,
    // coefficient 0). Keys [row*64 + l][4], -1 padded; W' coefficients as
    // (c, +-c) for (Re, Im) with the conjugation folded in.
    int wq_nblk, fq_nblk;
    const int* wq_boff;        // [wq_nblk+1] first term row of each block, times 64
    const int* wq_item;        // [wq_nblk*64] nonzero of each lane, -1 past the end
    const int* wq_keys;
    const double* wq_cc;
    const float* wq_ccf;
    const int* fq_boff;
    const int* fq_item;        // feature of each lane
    const int* fq_keys;
    const double* fq_c;
    const float* fq_cf;
    const int* lin_global;     // [n_lin] global linear feature id
};

struct ModelDev {
    int n_type, n_fn, n_tp, max_l, n_lmh;
    double cutoff;
    int n_features, n_linear;
    int tp_of[kMaxTypes * kMaxTypes];
    const double* rad;         // [n_tp * n_fn * 2] (beta, mu) by (tp, n_id)
    const int* rad_count;      // [n_tp] active radial functions
    TypeDev t[kMaxTypes];
};

// Column recipes of X: column c = prod_{k < nfac} x_{base_k}.
struct ColumnsDev {
    int n_cols;
    const unsigned char* nfac;        // [F]
    const int* base_local;            // [n_type][F][3] local feature ids or -1 (invalid)
    // column blocks
    int n_blocks;
    const int* blk_c0;                // [n_blocks+1] column ranges
    const int* need_off;              // [n_type][n_blocks+1]
    const int* need_feat;             // local feature ids needed by the block
    const int* col_need;              // [n_type][F][3] index into the block's need list
    const unsigned char* blk_kind;    // [n_blocks] 1: every column is a linear feature
    const int* blk_nkey;              // [n_blocks] Y keys the block's nonzeros use
};

// Polynomial columns built by the factor-derivative GEMM (kernels_poly.hip).
struct PolyDev {
    int K;                   // factors (distinct linear features of the products)
    int U;                   // monomials: K of degree 1, then K(K+1)/2 of degree 2
    int n_cols;              // polynomial columns
    int c0;                  // first polynomial column (= number of linear features)
    const int* fac_local;    // [n_type][K] local linear feature of each factor, -1 if none
    const int* col_phi;      // [n_cols][3] factor of each slot (-1 unused)
    const int* col_u;        // [n_cols][3] monomial of the other factors of each slot
    const int* mono_a;       // [U] first factor of each monomial
    const int* mono_b;       // [U] second factor (-1 for degree 1)
    // entries: the distinct (factor, monomial) pairs of the columns' slots,
    // sorted by factor, so that the entries of one factor are a contiguous range
    int E;                   // number of entries
    const int* ent_u;        // [E] monomial of each entry
    const int* col_pos;      // [n_cols][3] entry of each slot of a column, -1 if none
    const int4* col_pos4;    // [n_cols] the same, one 16-byte record per column (w = -1)
    // tiles of the entry products: (factor, first entry, end entry), at most
    // 128 entries of one factor each
    int n_ztiles;
    const int* ztiles;       // [n_ztiles][3]
};

// Signature groups of the linear features: features of one atom type whose
// derivative weights W' use the same Y keys for every type pair (they differ
// in the radial index and the coefficients). A task is up to kGroupSlots such
// features (slots); tasks are sorted by atom type. The grouped positions of a
// task and type pair run [key][slot]; wg_src maps each position to the atom's
// W' nonzero stored there in the feature-major W' of the type (-1: none).
constexpr int kGroupSlots = 16;
struct GroupDev {
    int n_tasks;
    int ntp;
    const int* task_type;   // [n_tasks]
    const int* task_gcol;   // [n_tasks][16] global column of each slot, -1 when empty
    const int* tt_nkey;     // [n_tasks][ntp] Y keys of the task for the type pair
    const int* tt_keyoff;   // [n_tasks][ntp] offset into keys
    const int* keys;        // Y keys (-1: the constant Y_00 folded into W')
    const int* tt_woff;     // [n_tasks][ntp] first grouped position of the task and type pair
    const int* tt_nid;      // [n_tasks][ntp][16] radial index of each slot, -1 when empty
    const int* g2l;         // [n_type][n_lin] local linear feature of a global column, -1 if none
    const int* wg_src;      // [n_type][wg_stride] W' nonzero at each grouped position, -1: none
    int wg_stride;          // grouped positions per type (max over types)
};

// A batch of structures resident in device memory.
struct BatchDev {
    int n_st;
    int n_atoms_total;
    const int* st_atom_off;   // [n_st+1]
    const int* st_row_off;    // [n_st+1] row offsets of X (7 + 3N per structure)
    const int* types;         // [A]
    const int* atom_st;       // [A] structure of each atom
    const double* pos;        // [A*3] (possibly rewrapped) Cartesian positions
    const int* nb_off;        // [A+1] full neighbour list
    const int* nb_j;          // [M] global atom index
    const double* nb_d;       // [M*3] r_j - r_i + T
};

}  // namespace plm
