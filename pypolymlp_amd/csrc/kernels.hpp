// Kernel launchers (host-callable) of the HIP hot path.
#pragma once

#include <hip/hip_runtime.h>

#include "device_tables.hpp"

namespace plm {

struct XRowsParams;
struct BlockWz;

void launch_neighbor_full(hipStream_t st, int n, const int* atom_st, const int* st_atom_off,
                          const double* pos, const int* tr_off, const double* trans, double cut2,
                          double tol2, int* count, const int* nb_off, int* nb_j, int* nb_t,
                          double* nb_d);
void launch_neighbor_half(hipStream_t st, int n, const int* atom_st, const int* st_atom_off,
                          const double* pos, const int* tr_off, const double* trans, double cut2,
                          double tol2, double tol, int* count, const int* nb_off, int* nb_j,
                          int* nb_t, double* nb_d);

struct BinGridHost {
    int nb[3];
    int mexp[3];
    double inv[9];
    const int* tindex;  // device pointer
};
void launch_bin_assign(hipStream_t st, int n, const double* pos, const BinGridHost& g, int* bin,
                       int* wrap, int* bin_count);
void launch_bin_scatter(hipStream_t st, int n, const int* bin, const int* start, int* cursor,
                        int* sorted);
void launch_bin_pairs(hipStream_t st, int n, int half, const double* pos, const int* bin,
                      const int* wrap, const int* start, const int* sorted, const BinGridHost& g,
                      const double* trans, double cut2, double tol2, double tol, int* count,
                      const int* nb_off, int* nb_j, int* nb_t, double* nb_d);

size_t anlm_smem(const ModelDev& md, int elem, int nthreads, int max_noconj);
void launch_anlm(hipStream_t st, bool fp64, const ModelDev& md, const BatchDev& bd, void* a_full,
                 int a_stride, int atom0, int n_atoms, int max_noconj, int negate = 0);

struct XRowsParams {
    int CB;        // columns per block (lanes per pair): 64, 32 or 16
    int NW;        // waves per workgroup
    int need_max;  // needed features per block (max over types and blocks)
    int wz_max;    // derivative-weight nonzeros per block (max)
    int rows_max;  // 7 + 3N of the largest structure
    int ntp;       // type pairs
    int slot;      // elements of a wave's record slot (>= 64/CB * record prefix)
    int ld;        // leading dimension of X
    int st0, st_end;
    int row_base;  // first X row of the structure range
    int xcd_map;   // 1: all blocks of a structure on one XCD
    int gtile;     // 1: the row tile is the structure's block of X (rows beyond LDS)
    long pair_base;
};

// Per (type, column block) derivative-weight tables: the block's nonzeros
// (indices into the type's W' list), their packed (Y key + 1) << 8 | n_id,
// and per needed feature the nonzero range of each type pair.
struct BlockWz {
    const int* off;          // [n_type][n_blocks+1] into nz / nk
    const int* nz;           // type-level W' index of each nonzero
    const int* nk;           // packed key / radial index of each nonzero
    const int* zr;           // [need entries][n_tp+1] block-local nonzero offsets
};

struct PairRec;
size_t pairdata_smem(const ModelDev& md, int elem);
// tpo (optional): the records of each atom are then stored grouped by type
// pair (stable within a type pair), and tpo[(i - atom0) * (n_tp + 1) + tp] is
// the first pair index of type pair tp of atom i (the last entry its end).
void launch_pairdata(hipStream_t st, bool fp64, const ModelDev& md, const BatchDev& bd, int atom0,
                     int n_atoms, long pair_base, void* pd, const PairRec& R, int lmax,
                     int* tpo = nullptr);
// W'_i[g] for every nonzero g of the atom's type and the feature values
// x_f(i), for atoms [atom0, atom0 + n_atoms): w[(i - atom0) * w_stride + g],
// x[(i - atom0) * x_stride + f].
// Order parameters of the X path from the pair records (grouped by type pair,
// tpo from launch_pairdata): a_full[(i - atom0) * a_stride + k].
void launch_anlm_rec(hipStream_t st, bool fp64, const ModelDev& md, const BatchDev& bd, int atom0,
                     int n_atoms, long pair_base, const void* pd, const PairRec& R,
                     const int* tpo, void* a_full, int a_stride, int max_noconj);
void launch_wprime(hipStream_t st, bool fp64, const ModelDev& md, const BatchDev& bd,
                   const void* a_full, int a_stride, int atom0, int n_atoms, void* w, int w_stride,
                   void* x, int x_stride, int max_nlm, int w_maxf, int v_maxf);

// Linear columns by signature groups (kernels_xgroup.hip).
struct XGroupParams {
    int TB;         // tasks (column groups of NSW slots) per workgroup
    int A;          // waves per task: atom streams
    int vt0, vt1;   // virtual task range of the launch (one atom type)
    int n_blocks;   // task blocks per structure
    int ld;         // leading dimension of X
    int st0, st_end;
    int row_base;   // first X row of the structure range
    int xcd_map;    // 1: all blocks of a structure on one XCD (static grid)
    int atom_base;  // first atom of the work buffers
    long pair_base;
    int split;      // slot parts per task (1, 2, 4; > 1 needs TB = 1): the row
                    // tile of a part holds NSW / split slot columns, so that the
                    // rows of large structures fit LDS
    int* queue;     // persistent launch: 8 item counters, zero before the
                    // launch (one queue per XCD); null: one workgroup per item
    int grid;       // persistent launch: workgroups
    int w_stride;   // complex W' entries per atom (launch_wprime's w_stride)
};
int xgroup_slots(bool fp64);      // slots per wave (NSW): 16 in FP32, 8 in FP64
int xgroup_max_waves(bool fp64);  // waves per workgroup the kernel is built for
size_t xgroup_smem(const XGroupParams& P, int rows_max, bool fp64);
// resident workgroups of a persistent launch with P's shape (all CUs)
int xgroup_persistent_grid(const XGroupParams& P, int rows_max, bool fp64);
// w: the atoms' feature-major W' (launch_wprime), read through gd.wg_src
void launch_xgroup(hipStream_t st, bool fp64, const GroupDev& gd, const BatchDev& bd,
                   const XGroupParams& P, int rows_max, const void* w, const void* pd,
                   const PairRec& R, const int* tpo, void* X, const double* roww);
// energy rows of the linear columns [0, n_lin) of structures [st0, st0 + n_st)
// from the feature values x (wprime), with the scale sums of the raw rows
void launch_xenergy(hipStream_t st, bool fp64, const GroupDev& gd, const BatchDev& bd, int n_lin,
                    int st0, int n_st, int row_base, int atom_base, const void* x, int x_stride,
                    void* X, int ld, const double* roww, double* xe_sum, double* xe_sq);
size_t xrows_smem(const XRowsParams& P, int elem);
int xrows_stage_limit(bool fp64);
int xrows_max_waves(bool fp64);  // waves per workgroup the kernel is built for
void launch_xrows(hipStream_t st, bool fp64, const ModelDev& md, const ColumnsDev& cd,
                  const BatchDev& bd, const BlockWz& bw, const XRowsParams& P, const void* w,
                  int w_stride, const void* x, int x_stride, int atom_base, const void* pd,
                  const PairRec& R, void* X, const double* roww, double* xe_sum, double* xe_sq);

struct EvalDev;
void launch_prodsum(hipStream_t st, bool fp64, const ModelDev& md, const EvalDev& ed,
                    const BatchDev& bd, const void* a_full, int a_stride, void* ps_e, void* ps_f,
                    int ps_stride, int max_lin, int pair_model);
// x[0, n) = 0 by a kernel (a graph of it replays correctly; a memset node is avoided)
void launch_fill_zero(hipStream_t st, double* x, size_t n);
// half-pair loop over the n_half pairs of the half list (hb_off[n_atoms] = n_half)
void launch_pair(hipStream_t st, bool fp64, const ModelDev& md, const EvalDev& ed,
                 const BatchDev& bd, const int* atom_st, const int* hb_off, const int* hb_j,
                 const double* hb_d, int n_half, const void* ps_e, const void* ps_f, int ps_stride,
                 double* energy, double* forces, double* stress, int pair_model);

// X^T X with the weighted targets in column F of X (ld >= F + 1, multiple of
// 128; rows_pad multiple of 64). Mixed: quantise X into the int8 slice buffer
// (gram_quant_bytes) then accumulate the exact Gram of the quantised X.
void launch_set_column(hipStream_t st, bool fp64, void* X, int ld, int col, const double* y,
                       int rows);
void launch_gram_f64(hipStream_t st, const double* X, int ld, int rows_pad, int F_pad, double* C,
                     int ldc);
// Row q of the quantised matrix is row rmap[q] of X (zero where rmap[q] < 0).
// Quantisation exponents are shared per column over chunks of rows: the first
// h0 rows (a multiple of 64, at most gram_chunk_rows()), then
// gram_chunk_rows() rows each.
int gram_chunk_rows();
size_t gram_quant_bytes(int rows_pad, int h0, int ld);
void launch_gram_quantize(hipStream_t st, bool x_fp64, const void* X, const int* rmap, int ld,
                          int rows_pad, int h0, int F_pad, void* quant);
void launch_gram_i8(hipStream_t st, int ld, int rows_pad, int h0, int F_pad, const void* quant,
                    double* C, int ldc);

// Residual pass (kernels_resid.hip): U[r][k] = X[r][F] - X[r][:ld] C[k][:ld]
// (C rows zero from column F on), R[k][:ld] += X^T U[:, k], S[k] += U[:, k]^T U[:, k],
// for n (1 to 8) coefficient rows C[k] of length ld; U holds rows x n.
void launch_residual_pass(hipStream_t st, bool fp64, const void* X, int ld, int rows, int F,
                          const double* C, int n, double* U, double* R, double* S);

// Polynomial columns (kernels_poly.hip)
void launch_poly_factor(hipStream_t st, bool fp64, const ModelDev& md, const PolyDev& pd,
                        const BatchDev& bd, int atom0, int n_atoms, const void* xall, int x_stride,
                        void* XF, void* M);
void launch_poly_deriv(hipStream_t st, bool fp64, const ModelDev& md, const PolyDev& pd,
                       const BatchDev& bd, int atom0, int n_atoms, long pair_base,
                       const void* prec, const PairRec& R, const void* wall, int w_stride, int k0,
                       int kc, int rd_max, void* D, int st0, const long* d_off);
// the polynomial columns of the n_st equal-size structures from s_first (rd
// rows each) of the sub-batch starting at structure s0 / atom atom_base
// (kernels_polyz.hip): the entry products Z, then the force / stress rows from
// Z and the energy rows from the factor values XF
void launch_poly_z(hipStream_t st, bool fp64, const PolyDev& pd, const BatchDev& bd, int s_first,
                   int n_st, int s0, int atom_base, int rd, const void* D, const long* d_off,
                   const void* M, void* Z, long z_stride);
void launch_poly_rows(hipStream_t st, bool fp64, const PolyDev& pd, const BatchDev& bd,
                      int s_first, int n_st, int atom_base, int rd, const void* XF, const void* Z,
                      long z_stride, void* X, int ld, int row_base, const double* roww,
                      double* xe_sum, double* xe_sq);


// Surrogate network of the hyperparameter search (FP32 MLP 5-64-64-2, SiLU).
constexpr int kSurrIn = 5, kSurrHidden = 64, kSurrOut = 2;
constexpr int kSurrParams = kSurrHidden * kSurrIn + kSurrHidden + kSurrHidden * kSurrHidden +
                            kSurrHidden + kSurrOut * kSurrHidden + kSurrOut;
constexpr int kSurrMaxBatch = 128;
size_t surrogate_smem(int n);
void launch_surr_train(hipStream_t st, float* P, float* M, float* V, const float* X,
                       const float* T, int n, int steps, int t0, float lr, float beta1,
                       float beta2, float eps, float wd, int mode, float* grad, float* loss);
void launch_surr_predict(hipStream_t st, const float* P, const float* X, int n, float* Y);

}  // namespace plm
