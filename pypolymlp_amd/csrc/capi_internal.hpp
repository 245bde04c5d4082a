// Internal handles behind the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <tuple>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "device_tables.hpp"
#include "eval_tables.hpp"
#include "kernels.hpp"
#include "pair_data.hpp"
#include "model.hpp"

namespace plm {

extern thread_local std::string g_last_error;

struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void check_hip(hipError_t e, const char* what);

struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf();
    void alloc(size_t n);
    void release();
};

struct Device {
    bool init = false;
    hipStream_t stream = nullptr;
};
Device& device();

// Owns many small device arrays uploaded once.
struct Store {
    std::vector<std::unique_ptr<DevBuf>> bufs;
    template <class T>
    const T* upload(const T* host, size_t bytes) {
        auto b = std::make_unique<DevBuf>();
        b->alloc(bytes ? bytes : 16);
        if (bytes) check_hip(hipMemcpy(b->ptr, host, bytes, hipMemcpyHostToDevice), "upload");
        const T* p = reinterpret_cast<const T*>(b->ptr);
        bufs.push_back(std::move(b));
        return p;
    }
};

struct ColumnTables {
    Store store;
    ColumnsDev cd{};
    BlockWz bw{};
    int need_max = 0, wz_max = 0, CB = 0, nkey_max = 0;
};

struct PolyTables {
    Store store;
    PolyDev dev{};
    bool enabled = false;
    // entries of factor f: [ent_off[f], ent_off[f + 1])
    std::vector<int> ent_off;
};

struct GroupTables {
    Store store;
    GroupDev dev{};
    bool enabled = false;
    std::vector<int> type_task0;  // [n_type + 1] first task of each atom type
};

struct ModelHandle {
    Model host;
    Store store;
    ModelDev md{};
    bool built = false;
    int max_nlm = 0, max_noconj = 0;
    std::vector<std::vector<int>> host_wz_off;
    std::vector<std::vector<int>> host_wz_ylm;  // Y key of each derivative-weight nonzero
    std::vector<std::vector<int>> host_wz_nid;  // radial index of each nonzero
    int wz_max_type = 0, lin_max_type = 0;      // nonzeros / linear features, max over types
    int w_maxf = 0, v_maxf = 0;  // factors per W' term / per value term, max over types
    std::map<int, std::unique_ptr<ColumnTables>> col_cache;
    // evaluation tables
    Store pot_store;
    bool pot_uploaded = false;
    unsigned pot_version = 0;  // counts uploads of the potential tables
    EvalDev ed{};
    int max_heads = 0, max_lin = 0;
    // per type, the position of each head (in the potential's order) in the
    // per-atom head sums, which follow the half-pair loop's order
    std::vector<std::vector<int>> head_rank;
    explicit ModelHandle(const ModelInput& in);
    void build_device();
    // column-block tables for blocks of CB columns over columns [0, cover)
    const ColumnTables& columns(int CB, int cover);
    // tables of the polynomial columns built by the factor-derivative GEMM
    std::unique_ptr<PolyTables> poly_tables;
    const PolyTables& poly();
    // signature groups of the linear features (kernels_xgroup.hip)
    std::unique_ptr<GroupTables> group_tables;
    const GroupTables& groups();
    // launch shapes of the signature-group kernel chosen by timing, per
    // (precision, row bucket of the largest structure, structure-count bucket)
    struct GroupChoice {
        int TB = 0, A = 0, split = 1;
        bool persistent = false;
        double ms = 0.0;       // time of the chosen shape on the tuning structures
        double rule_ms = 0.0;  // time of the rule-based shape on the same structures
        int candidates = 0;
    };
    std::map<std::tuple<int, int, int>, GroupChoice> group_tuning;
    // work buffers of eager evaluations (order parameters, head sums, results),
    // kept and grown across calls; eval_mutex serialises the calls that use them
    DevBuf ev_a, ev_pse, ev_psf, ev_out;
    std::mutex eval_mutex;
};

struct BatchHandle {
    ModelHandle* model;  // may be null for neighbour-list-only use
    double cutoff = 0.0;
    int n_st = 0, n_atoms = 0, max_atoms = 0, max_pairs_atom = 0;
    long n_pairs = 0;
    std::vector<int> atom_off, row_off, types, nb_off;
    DevBuf d_types, d_pos, d_trans, d_tr_off, d_atom_st, d_atom_off, d_row_off;
    DevBuf d_nb_off, d_nb_j, d_nb_t, d_nb_d;
    // half neighbour list (evaluation), built on first use
    bool has_half = false;
    long n_half = 0;
    DevBuf d_hb_off, d_hb_j, d_hb_t, d_hb_d;
    // cell-list search (single large structure)
    bool binned = false;
    BinGridHost grid{};
    unsigned long long serial = 0;  // creation number (identifies the batch's rows)
    DevBuf d_tindex, d_bin, d_wrap, d_bin_start, d_bin_sorted;
    BatchHandle(ModelHandle* m, double cutoff, int n_st, const double* axes, const double* pos,
                const int* types, const int* n_atoms);
    BatchDev view() const;
    void ensure_half();
    // neighbour list (full or half) into the given buffers; returns pair count
    long build_list(bool half, DevBuf& off, DevBuf& nj, DevBuf& nt, DevBuf& nd,
                    std::vector<int>& host_off);
    void prepare_bins();
};

struct XPlan {
    XRowsParams P;
    const ColumnTables* ct;
    size_t smem;
};
XPlan plan_x(ModelHandle& m, int rows_max, bool fp64, int cover);

// Work buffers of the X build (grown as needed): order parameters, W' and
// feature values per atom, pair records.
struct XWork {
    DevBuf a, w, x, pd;
    DevBuf tpo;                // per-atom type-pair ranges of the pair records
    DevBuf queue;              // work-queue counters of the persistent signature-group launches
    DevBuf tune_x;             // scratch rows for timing signature-group launch shapes
    DevBuf xf, m, d, doff, z;  // polynomial columns: factor values, monomials, D, entry products
};

// X rows of structures [s0, s1) into X (row-major, leading dimension ld),
// weighted by roww when given.
// x_is_zero: X's rows of the range are already zero.
// poly_doff: device D offsets of the sub-batch (poly_d_offsets) or null
void build_x_rows(ModelHandle& m, BatchHandle& b, int s0, int s1, bool fp64, void* X, int ld,
                  const double* roww, double* xe_sum, double* xe_sq, XWork& work,
                  bool x_is_zero = false, const long* poly_doff = nullptr);
// device bytes per atom of the X build's per-atom work buffers
size_t atom_work_bytes(ModelHandle& m, bool fp64);
// polynomial columns of structures [s0, s1) (after the linear columns of the
// same sub-batch have been built, which leaves the W', x and pair buffers)
void build_poly_rows(ModelHandle& m, BatchHandle& b, int s0, int s1, bool fp64, void* X, int ld,
                     const double* roww, double* xe_sum, double* xe_sq, XWork& work,
                     long pair_base, const PairRec& R, int w_stride, int x_stride,
                     const long* doff_dev = nullptr);
// offsets of the factor derivatives D of structures [s0, s1), from 0
std::vector<long> poly_d_offsets(ModelHandle& m, const BatchHandle& b, int s0, int s1);
size_t poly_bytes(ModelHandle& m, const BatchHandle& b, int s0, int s1, bool fp64);
size_t pair_record_bytes(const ModelHandle& m, bool fp64);

// Work buffers of an accumulation: X of a sub-batch, its quantised slices,
// and the X build's per-atom and per-pair buffers (grown as needed).
struct AccumWork {
    DevBuf xbuf, qbuf;
    XWork work;
};

// The weighted rows [X | y] of one sub-batch kept in device memory after
// the accumulation, for a later residual pass over the same rows.
struct KeptRows {
    unsigned long long batch = 0, wy_hash = 0;  // batch serial, hash of the weights and targets
    int s0 = 0, s1 = 0;
    bool fp64 = false;
    DevBuf x;
};

// Cholesky factors of an accumulator's scaled system, kept by plm_normal_ridge
// for plm_normal_ridge_correct: valid for the accumulator's state (rows
// accumulated) and the scales / dropped columns they were made with.
struct FactorCache {
    long n_rows = -1;
    std::vector<double> scales;
    std::vector<unsigned char> zero;
    std::vector<std::pair<double, std::unique_ptr<DevBuf>>> f;  // (alpha, factor)
    size_t bytes = 0;
    void clear() {
        f.clear();
        bytes = 0;
        n_rows = -1;
    }
};

struct NormalHandle {
    ModelHandle* model = nullptr;
    int F = 0, F_pad = 0;
    long n_rows = 0;
    DevBuf xtx;   // F_pad x F_pad Gram of [X | y], upper triangle accumulated
    DevBuf vec;   // [xe_sum | xe_sq_sum], each F_pad
    AccumWork ws;
    size_t keep_budget = 0, kept_bytes = 0;  // plm_normal_keep_rows
    std::vector<std::unique_ptr<KeptRows>> kept;
    FactorCache factors;
};
// FNV-1a over the row weights and weighted targets of n rows
unsigned long long hash_rows(const double* w, const double* y, long n);

// fp64: the X build in FP64 (else FP32); quant: the exact three-slice int8
// Gram (else the FP64 Gram, which needs fp64)
void normal_accumulate(NormalHandle& acc, BatchHandle& b, const double* row_w, const double* y_w,
                       bool fp64, bool quant);

// The accumulation of one batch as a host-side plan (sub-batches, row maps,
// D offsets, weights and targets uploaded once) and a launch-only sequence,
// so that it can run eagerly or be captured as a HIP graph.
struct SubBatch {
    int s0, s1;
    long rows, rows_pad, nes_pad, h0, rows_q;  // rows_q / h0: quantised rows, first chunk
    long rmap_off, doff_off;                    // into AccumPlan::rmaps / doffs
};
struct AccumPlan {
    bool fp64 = true, quant = false;
    long n_rows = 0;
    std::vector<SubBatch> sb;
    DevBuf rmaps, doffs, w, y;
};
// quant: also the quantisation row maps (the int8 Gram's rows)
void plan_accumulate(NormalHandle& acc, BatchHandle& b, const double* row_w, const double* y_w,
                     bool fp64, bool quant, AccumPlan& plan);
// launches only (no allocation beyond growing ws, no copy, no synchronisation);
// with keep (eager accumulation only) each sub-batch's rows go to a buffer of
// their own, kept on the accumulator while its budget allows
struct KeepRequest {
    unsigned long long batch, wy_hash;
};
void enqueue_accumulate(NormalHandle& acc, BatchHandle& b, const AccumPlan& plan, AccumWork& ws,
                        double* gram, double* vec, hipStream_t st,
                        const KeepRequest* keep = nullptr);

// The fit step (one accumulation pass over a batch) captured as a HIP graph.
struct NormalGraph {
    NormalHandle* acc = nullptr;
    BatchHandle* b = nullptr;
    AccumPlan plan;
    AccumWork ws;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    NormalGraph() = default;
    NormalGraph(const NormalGraph&) = delete;
    NormalGraph& operator=(const NormalGraph&) = delete;
    ~NormalGraph();
};
void normal_graph_capture(NormalGraph& g, const double* row_w, const double* y_w, bool fp64,
                          bool quant);
void normal_graph_run(NormalGraph& g, int n_steps);

// Residual pass over a batch's weighted rows for n_coef (1..8) unscaled
// coefficient vectors (rows of coeffs, n_coef x F): xtr[k] += X^T (y - X c_k)
// and sq[k] += |y - X c_k|^2, with X built in FP64 or FP32. The accumulator
// supplies the model and its work buffers; its sums are not touched.
void normal_residual(NormalHandle& acc, BatchHandle& b, const double* row_w, const double* y_w,
                     bool fp64, const double* coeffs, int n_coef, double* xtr, double* sq);
// keep the rows of later eager accumulations up to max_bytes (0: keep none,
// release the kept rows)
void normal_keep_rows(NormalHandle& acc, size_t max_bytes);

void upload_potential(ModelHandle& m);

// One evaluation of a device-resident batch (zero fill, order parameters,
// per-atom contraction, half-pair loop) captured as a HIP graph and replayed.
// The graph holds the model's potential tables and the batch's neighbour
// lists as of capture; run() refuses a model whose coefficients changed since.
struct EvalGraph {
    ModelHandle* m = nullptr;
    BatchHandle* b = nullptr;
    bool fp64 = true;
    unsigned pot_version = 0;
    size_t n_out = 0;
    DevBuf a_buf, pse, psf, out;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    EvalGraph() = default;
    EvalGraph(const EvalGraph&) = delete;
    EvalGraph& operator=(const EvalGraph&) = delete;
    ~EvalGraph();
};
void enqueue_eval_dev(ModelHandle& m, const BatchDev& bd, const int* hb_off, const int* hb_j,
                      const double* hb_d, int n_half, bool fp64, void* a_buf, void* pse,
                      void* psf, double* out, size_t n_out, hipStream_t st);
void eval_graph_capture(EvalGraph& g);
void eval_graph_run(EvalGraph& g, int steps, double* energies, double* forces, double* stresses);
void eval_single(ModelHandle& m, const double* axis, const double* pos_c, const int* types, int n,
                 bool fp64, double* energy, double* forces, double* stress);
// plm_eval with every device buffer in a fixed file-scope device array
// (eval_noalloc.hip); refuses structures whose buffers exceed it
void eval_single_noalloc(ModelHandle& m, const double* axis, const double* pos_c, const int* types,
                         int n, bool fp64, double* energy, double* forces, double* stress);
void eval_batch(ModelHandle& m, BatchHandle& b, bool fp64, double* energies, double* forces,
                double* stresses);
void ridge_solve(int F, const double* xtx, const double* xty, const double* alphas, int n_alpha,
                 double* coeffs, int* info);
// ridge / squared errors on an accumulator's own normal equations, scaled on
// the device (no read-back of X^T X)
void normal_ridge(NormalHandle& acc, const double* scales, const unsigned char* zero,
                  const double* alphas, int n_alpha, double* coeffs, int* info, double* sq_err);
void normal_sq_error(NormalHandle& acc, const double* scales, const unsigned char* zero,
                     const double* coeffs, int n_coef, double* sq_err);
void normal_ridge_correct(NormalHandle& acc, const double* scales, const unsigned char* zero,
                          const double* alphas, int n_alpha, const double* rhs, double* delta,
                          int* info, double* dad);
// CG on (xtx + alpha I) c = xty from c0 (zeros when null); returns the
// iteration at which the stopping test ended the loop, or max_iter
int cg_solve(int F, const double* xtx, const double* xty, double alpha, const double* coef0,
             double gtol, int max_iter, double* coef);

void surrogate_run(float* params, float* adam_m, float* adam_v, int t0, const float* x,
                   const float* y, int n, int steps, float lr, float wd, int mode, float* grad,
                   float* loss);
void surrogate_predict(const float* params, const float* x, int n, float* y);

}  // namespace plm
