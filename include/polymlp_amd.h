/*
 * polymlp_amd: C-ABI of the MI355X (gfx950) hot path of pypolymlp.
 *
 * Plain pointers and sizes only. Every function returns 0 on success and a
 * non-zero code on failure; plm_last_error() then describes the failure
 * (thread-local). Nothing here calls exit().
 *
 * Array conventions (as the reference's pybind11 module receives them):
 *   axis   : 3x3 row-major, lattice vectors as columns (axis[i][col])
 *   pos_c  : Cartesian positions, 3 x n row-major (component-major), i.e.
 *            axis @ fractional positions, not rewrapped
 *   types  : n int32 atom types (0-based, model element order)
 * Batches concatenate these per structure: axes (n_st*9), pos (sum 3*n_s,
 * each structure's 3 x n_s block contiguous), types (sum n_s), n_atoms (n_st).
 */
#ifndef POLYMLP_AMD_H
#define POLYMLP_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define PLM_OK 0
#define PLM_ERR_ARG 1
#define PLM_ERR_MODEL 2
#define PLM_ERR_DEVICE 3
#define PLM_ERR_MEMORY 4
#define PLM_ERR_SOLVE 5

/* Precision of the feature stream / Gram accumulation. */
#define PLM_FP64 0  /* plain FP64 features, FP64 MFMA X^T X (the drop-in form) */
#define PLM_MIXED 1 /* FP32 features; X^T X as the exact Gram of X quantised per
                     * (row chunk, column) to a shared exponent and three signed
                     * 7-bit slices, on int8 MFMA with int32 sums combined in FP64 */
#define PLM_MIXED64 2 /* FP64 features with the quantised int8 Gram of PLM_MIXED
                       * (X^T X accumulation; elsewhere the same as PLM_FP64) */

typedef struct plm_model plm_model;
typedef struct plm_batch plm_batch;
typedef struct plm_normal plm_normal;
typedef struct plm_eval_graph plm_eval_graph;
typedef struct plm_normal_graph plm_normal_graph;

/* Model description: the contents of the params dict the reference's
 * convert_params_dict_to_feature_params reads
 * (src/cxx/src/compute/py_params.cpp:11-60). */
typedef struct plm_model_desc {
    int n_type;
    int feature_type;     /* 0 = pair, 1 = gtinv */
    int model_type;       /* 1..4 */
    int max_p;            /* 1..3 */
    int max_l;
    double cutoff;
    int n_params;
    const double* params; /* n_params x 2: (beta, mu) */
    /* pair_params_conditional for type pairs (i <= j) in the order
     * (0,0),(0,1),...,(0,n-1),(1,1),...: counts then concatenated indices */
    const int* cond_count;
    const int* cond_index;
    /* invariant tables (Readgtinv output, lm = l*l + l + m) */
    int n_lcomb;
    const int* lcomb_order;   /* n_lcomb */
    const int* lcomb_l;       /* sum(order) */
    const int* lcomb_nterm;   /* n_lcomb */
    const int* lm;            /* sum(order_i * nterm_i) */
    const double* lm_coeff;   /* sum(nterm) */
} plm_model_desc;

const char* plm_last_error(void);
int plm_device_count(int* n);

/* Model setup: Mapping / ModelParams / Features / FeaturesPoly
 * (src/cxx/src/polymlp/polymlp_features.cpp:27-59). Uploads the tables. */
int plm_model_create(const plm_model_desc* desc, plm_model** out);
void plm_model_destroy(plm_model* m);
/* PolymlpAPI::get_n_variables (polymlp_api.cpp:190-197) */
int plm_model_n_features(const plm_model* m, int* n);
/* Per-type sizes: order-parameter components, non-conjugate ones, linear features. */
int plm_model_type_sizes(const plm_model* m, int type, int* n_nlm, int* n_noconj, int* n_linear);
/* Potential fold for evaluation (PolymlpEval ctor, polymlp_eval.cpp:12-20;
 * Potential, polymlp_potential.cpp:25-156). coeffs = coeffs / scales. */
int plm_model_set_coeffs(plm_model* m, const double* coeffs, int n);

/* NeighborCell (src/cxx/src/compute/neighbor_cell.cpp:11-257): lattice
 * translations (n_trans x 3) and the positions used for the search (3 x n,
 * rewrapped only for reduced cells). Two-call: trans == NULL returns n_trans. */
int plm_cell_translations(const double* axis, const double* pos_c, int n, double cutoff,
                          double* trans, int max_trans, double* pos_rev, int* n_trans);

/* NeighborFull / NeighborHalf (neighbor_full.cpp:10-76, neighbor_half.cpp:10-100)
 * on the GPU. offset (n+1), neigh (M), trans_id (M), d (M x 3, r_j - r_i + T).
 * Two-call: with capacity < M only n_pairs is set. */
int plm_neighbor(const double* axis, const double* pos_c, int n, double cutoff, int half,
                 int* offset, int* neigh, int* trans_id, double* d, long capacity, long* n_pairs);

/* Device-resident batch of structures (positions, translations, full
 * neighbour list). */
int plm_batch_create(const plm_model* m, int n_st, const double* axes, const double* pos,
                     const int* types, const int* n_atoms, plm_batch** out);
void plm_batch_destroy(plm_batch* b);
/* rows of X for the batch with forces and stresses: sum(7 + 3 n_s) */
int plm_batch_n_rows(const plm_batch* b, long* rows);
/* Neighbour pairs (full list) of the batch; no reference counterpart (batch
 * statistics for throughput accounting). */
int plm_batch_n_pairs(const plm_batch* b, long* pairs);

/* PotentialModel.get_x (py_model.cpp:10-54): X (rows x F, row-major) with the
 * reference's row layout [n_st energy rows | 6 stress rows per structure |
 * 3N force rows per structure], all structures with forces. precision =
 * PLM_FP64 or PLM_MIXED (FP32 stream, written to FP64). */
int plm_features_x(plm_model* m, plm_batch* b, int precision, double* x_out);

/* Fused X + X^T X accumulation (data_sequential.py:25-155 with sum_xtx,
 * utils_sequential.py:94-109). row_w / row_y: per-row weights and weighted
 * targets in the batch's row order (plm_batch_row_order). X^T X accumulates
 * (w x)(w x)^T over rows; X^T y accumulates (w x) * y; the raw (unweighted)
 * energy rows feed xe_sum and xe_sq_sum as in the reference. */
int plm_normal_create(const plm_model* m, plm_normal** out);
void plm_normal_destroy(plm_normal* acc);
int plm_normal_accumulate(plm_normal* acc, plm_batch* b, const double* row_w, const double* y_w,
                          int precision);
/* xtx: F x F full symmetric (row-major); xty, xe_sum, xe_sq_sum: F;
 * y_sq: sum of squared weighted targets. X^T y and y^T y are accumulated by
 * the same Gram pass as X^T X (targets as column F of X), so all three come
 * from the same X; in PLM_MIXED that X is the FP32 X quantised to three
 * 7-bit slices per (4096-row chunk, column) exponent, whose Gram is exact.
 * Any output may be NULL. */
int plm_normal_result(plm_normal* acc, double* xtx, double* xty, double* xe_sum,
                      double* xe_sq_sum, double* y_sq, long* n_rows);
/* The fit step captured as a HIP graph: one plm_normal_accumulate pass of a
 * device-resident batch with fixed row weights and targets
 * (_compute_products_single_batch, data_sequential.py:97-155, over the
 * batch's sub-batches: X rows, quantisation and Gram) recorded once and
 * replayed. Each replay adds the batch's contribution to the accumulator,
 * exactly as another plm_normal_accumulate call would. Creation runs one
 * eager pass into scratch sums (the accumulator is unchanged). The graph
 * holds its own work buffers; the accumulator and batch must outlive it. */
int plm_normal_graph_create(plm_normal* acc, plm_batch* b, const double* row_w, const double* y_w,
                            int precision, plm_normal_graph** out);
int plm_normal_graph_run(plm_normal_graph* g, int n_steps);
int plm_normal_graph_destroy(plm_normal_graph* g);

/* Row permutation: for each row of the batch in the reference layout, the
 * batch-internal row index (batch rows are [e, 6 s, 3N f] per structure). */
int plm_batch_row_order(const plm_batch* b, long* order);

/* Surrogate network of the hyperparameter search (no reference counterpart:
 * the reference fits every grid candidate; utils/grid_search is the
 * exhaustive fallback this network prunes). FP32 MLP 5-64-64-2 with SiLU,
 * parameters [W1 (64x5) | b1 | W2 (64x64) | b2 | W3 (2x64) | b3], loss =
 * mean (y - t)^2. Training runs `steps` full-batch Adam steps (beta 0.9 /
 * 0.999, eps 1e-8, decoupled weight decay; step count continues from t0) in
 * one launch; params and the Adam moments are updated in place. n <= 128 for
 * train / grad. */
int plm_surrogate_n_params(int* n);
int plm_surrogate_train(float* params, float* adam_m, float* adam_v, int t0, const float* x,
                        const float* y, int n, int steps, float lr, float weight_decay,
                        float* loss);
int plm_surrogate_grad(const float* params, const float* x, const float* y, int n, float* grad,
                       float* loss);
int plm_surrogate_predict(const float* params, const float* x, int n, float* y);

/* Ridge regression on the normal equations (solvers.py:9-84): for each alpha
 * (ascending), Cholesky of (xtx + alpha I) in FP64 on the GPU, solve, and one
 * step of iterative refinement. xtx: F x F symmetric; coeffs: n_alpha x F.
 * info[k] != 0 marks a failed factorization (the reference then returns
 * coefficients of 1e30). */
int plm_ridge(int F, const double* xtx, const double* xty, const double* alphas, int n_alpha,
              double* coeffs, int* info);

/* The ridge solve on an accumulator's own normal equations, without reading
 * X^T X back: the scaling of calc_xtx_xty (data_sequential.py:72-92: columns
 * with zero[f] != 0 dropped, X^T X / (s s^T), X^T y / s; zero may be NULL) is
 * applied on the device, then solver_ridge (solvers.py:48-84) runs as in
 * plm_ridge. coeffs: n_alpha x F (coefficients of the scaled problem), info:
 * n_alpha. sq_err (n_alpha, may be NULL): c^T A c - 2 c^T b + y^T y of each
 * solution on the same scaled system, the numerator of compute_rmse
 * (utils_model_selection.py:37-47). */
int plm_normal_ridge(plm_normal* acc, const double* scales, const unsigned char* zero,
                     const double* alphas, int n_alpha, double* coeffs, int* info,
                     double* sq_err);
/* compute_rmse's numerator (utils_model_selection.py:37-47) of n_coef
 * coefficient vectors (n_coef x F) against an accumulator's normal equations
 * scaled by scales / zero on the device (e.g. the test set with the training
 * scales). */
int plm_normal_sq_error(plm_normal* acc, const double* scales, const unsigned char* zero,
                        const double* coeffs, int n_coef, double* sq_err);

/* Residual pass over a batch for n_coef (1..8) unscaled coefficient vectors
 * (coeffs: n_coef x F): xtr[k] += X^T (y - X c_k) (n_coef x F) and
 * sq[k] += |y - X c_k|^2, from the same weighted rows [X | y] that
 * plm_normal_accumulate folds into the Gram (row_w / y_w in batch order; X
 * built in FP64 for PLM_FP64 and PLM_MIXED64, FP32 for PLM_MIXED). No Gram is
 * formed and the accumulator's sums are not changed; it supplies the model
 * and the work buffers. sq is compute_rmse's numerator
 * (utils_model_selection.py:37-47) taken directly from the rows; xtr is the
 * residual of the normal equations of X itself, for the refinement step of
 * plm_normal_ridge_correct. The reference solves in FP64 from an FP64 Gram
 * and needs no such step; this serves its fit driver (fit.py:12-63) when the
 * Gram is the quantised one. */
int plm_normal_residual(plm_normal* acc, plm_batch* b, const double* row_w, const double* y_w,
                        int precision, const double* coeffs, int n_coef, double* xtr, double* sq);
/* Keep the weighted rows [X | y] that later plm_normal_accumulate calls build
 * (per sub-batch, while their total stays within max_bytes of device
 * memory), so that plm_normal_residual over the same batch with the same
 * row_w / y_w and precision reads them instead of building them again.
 * max_bytes = 0 releases the kept rows (and the Cholesky factors that
 * plm_normal_ridge keeps for plm_normal_ridge_correct) and keeps none. No counterpart in the
 * reference (its X is a host array, data_sequential.py:97-155); it serves the
 * refinement step of a fit from the quantised Gram. */
int plm_normal_keep_rows(plm_normal* acc, unsigned long long max_bytes);
/* Correction step of iterative refinement on an accumulator's scaled system
 * (scaling as in plm_normal_ridge): delta_k = (A + alpha_k I)^{-1} rhs_k
 * (Cholesky of the accumulated Gram's system in FP64, with one inner
 * refinement step), rhs / delta: n_alpha x F, info as in plm_normal_ridge.
 * The factors of the last plm_normal_ridge call on the same accumulator state
 * and scaling are reused (plm_normal_ridge keeps up to 32 GiB of them until the
 * next accumulation).
 * With rhs_k = xtr_k / s - alpha_k c_k from plm_normal_residual, c_k + delta_k
 * solves the normal equations of X to the accuracy of the X build instead of
 * that of the quantised Gram. dad (n_alpha, may be NULL): delta_k^T A delta_k,
 * for the squared error after the step. No counterpart in the reference (its
 * Gram is FP64); serves its fit driver (fit.py:12-63). */
int plm_normal_ridge_correct(plm_normal* acc, const double* scales, const unsigned char* zero,
                             const double* alphas, int n_alpha, const double* rhs, double* delta,
                             int* info, double* dad);

/* The launch shapes the autotuner chose for this model so far, one text line
 * per class (precision, largest-structure row bucket, structure-count
 * bucket): tasks per workgroup, waves per task, slot parts, static grid or
 * persistent work queue, and its time against the rule-based shape. buf (may
 * be NULL) receives at most cap - 1 characters and a terminating zero; len (may
 * be NULL) the full length. No counterpart in the reference (its X build is
 * OpenMP over structures, py_model.cpp:41); it documents the shapes of this
 * library's signature-group kernel. */
int plm_model_tuning(const plm_model* m, char* buf, int cap, int* len);

/* FeaturesAttr (src/cxx/src/compute/py_features_attr.cpp:11-44): the
 * attributes of the columns of X. Per linear feature f (n_linear of them):
 * radial[f] its radial index, gtinv[f] its invariant (l-combination) index
 * (-1 for pair features), tcomb[6 f .. 6 f + 5] its type-pair combination
 * (-1 padded). Per polynomial column p (n_poly): poly[3 p ..] the linear
 * features of the product (-1 padded). type_pairs: n_type x n_type type-pair
 * index. Any output pointer may be NULL (call once with NULLs for the sizes). */
int plm_model_feature_attrs(const plm_model* m, int* n_linear, int* n_poly, int* radial,
                            int* gtinv, int* tcomb, int* poly, int* type_pairs);

/* Conjugate gradient (solver_cg, src/pypolymlp/mlp_dev/gradient/solvers_cg.py:32-93)
 * on (xtx + alpha I) coef = xty from coef0 (zeros when NULL), device resident:
 * the loop stops at the first iteration i > 100 with |r| / sqrt(F) < gtol, or
 * after max_iter iterations. coef: F; *n_iter (may be NULL) receives the
 * iteration at which the loop stopped. */
int plm_solver_cg(int F, const double* xtx, const double* xty, double alpha, const double* coef0,
                  double gtol, int max_iter, double* coef, int* n_iter);

/* PotentialPropertiesFast.eval (py_properties_fast.cpp:19-28): energy,
 * forces (n x 3) and stress (xx, yy, zz, xy, yz, zx) in eV/cell. */
int plm_eval(plm_model* m, const double* axis, const double* pos_c, const int* types, int n,
             int precision, double* energy, double* forces, double* stress);

/* No counterpart in the reference. plm_eval for callers that cannot allocate
 * device memory (serving loops, code inside a caller's stream capture,
 * harnesses that forbid hipMalloc): the same kernels and results as plm_eval,
 * with every device buffer (staged inputs, full and half neighbour lists,
 * order parameters, per-atom contractions, outputs) carved out of one
 * file-scope device array of 64 MiB per GPU, held until the process ends. It
 * calls no hipMalloc / hipFree. The model must already hold its coefficients
 * (plm_model_set_coeffs). A structure whose buffers exceed 64 MiB is refused
 * with PLM_ERR_ARG (about 6,000 atoms of a 1-type order-2 max_l 4 model).
 * Calls on one GPU share the array: overlapping calls on the same GPU race. */
int plm_eval_noalloc(plm_model* m, const double* axis, const double* pos_c, const int* types,
                     int n, int precision, double* energy, double* forces, double* stress);

/* Evaluation of a device-resident batch (eval_multiple, py_properties_fast.cpp:30-77). */
int plm_eval_batch(plm_model* m, plm_batch* b, int precision, double* energies, double* forces,
                   double* stresses);

/* Repeated evaluation of one device-resident batch, as in a molecular-dynamics
 * or relaxation loop that calls PolymlpEval::eval on the same cells
 * (src/cxx/src/polymlp_eval.cpp, reached from py_properties_fast.cpp:30-77):
 * the evaluation kernels of plm_eval_batch captured once as a HIP graph, then
 * replayed n_steps times per run call. The results are those of the last
 * replay. The graph keeps the model's coefficients and the batch's neighbour
 * lists as of creation: run fails with PLM_ERR_ARG after plm_model_set_coeffs,
 * and the model and batch must outlive the graph. */
int plm_eval_graph_create(plm_model* m, plm_batch* b, int precision, plm_eval_graph** out);
int plm_eval_graph_run(plm_eval_graph* g, int n_steps, double* energies, double* forces,
                       double* stresses);
int plm_eval_graph_destroy(plm_eval_graph* g);

/* libmlpcpp support functions on the device, running the device functions of
 * the X and evaluation kernels (PLM_FP64 or PLM_MIXED = FP32).
 * get_fn (src/cxx/src/python/pybind11_mlp.cpp:162-171,
 * polymlp_functions_interface.cpp:41-63): Gaussian x cosine-cutoff radial
 * functions at n distances r for n_params (beta, mu) pairs;
 * fn, fn_d: n x n_params. */
int plm_get_fn(int n, const double* r, const double* params, int n_params, double cutoff,
               int precision, double* fn, double* fn_d);
/* get_ylm (pybind11_mlp.cpp:173-184, polymlp_functions_interface.cpp:113-139):
 * Y_l^{-m} (m >= 0) at key l(l+3)/2 - m, l <= lmax, and their Cartesian
 * derivatives at n points xyz (n x 3, r = |xyz|). ylm: n x nh complex as
 * (re, im) pairs, dylm: n x nh x 3 (d/dx, d/dy, d/dz) complex pairs,
 * nh = (lmax + 1)(lmax + 2) / 2. */
int plm_get_ylm(int n, const double* xyz, int lmax, int precision, double* ylm, double* dylm);
/* PolymlpAPI::compute_anlmtp_conjugate + compute_features
 * (src/cxx/src/polymlp/polymlp_api.cpp:63-87, 130-140) for one atom of a gtinv
 * model's type: the non-conjugate order parameters (anlm_re, anlm_im, one per
 * non-conjugate component, plm_model_type_sizes) are completed with their
 * conjugates into anlm_full (n_nlm complex pairs) and the type's linear
 * feature values written to features (n_linear). Either output may be NULL. */
int plm_api_features(plm_model* m, int type, const double* anlm_re, const double* anlm_im,
                     int precision, double* anlm_full, double* features);
/* PolymlpAPI::compute_sum_of_prod_anlmtp (polymlp_api.cpp:174-185): prod_sum_e
 * and prod_sum_f (one complex pair per non-conjugate head) of the completed
 * order parameters, with the coefficients of plm_model_set_coeffs. */
int plm_api_sum_of_prod(plm_model* m, int type, const double* anlm_re, const double* anlm_im,
                        int precision, double* prod_e, double* prod_f);

/* Instrumentation (no counterpart in the reference): per-kernel-class timing
 * with HIP events recorded on the library's own stream, for roofline
 * figures. Names: "pairdata", "anlm", "wprime", "xenergy", "xgroup" (linear
 * columns by signature groups), "xrows" (linear columns by column blocks),
 * "poly_deriv", "poly_z", "poly_rows" (polynomial columns: factor
 * derivatives, entry products, column sums into the rows), "quantize",
 * "gram", "residual" (residual pass), "eval_anlm", "eval_prodsum",
 * "eval_pair", "cholesky", "chol_solve". */
int plm_profile_enable(int on);
int plm_profile_reset(void);
int plm_profile_get(const char* name, double* total_ms, long* launches);

#ifdef __cplusplus
}
#endif

#endif /* POLYMLP_AMD_H */
