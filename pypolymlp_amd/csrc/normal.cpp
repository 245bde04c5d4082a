// Fused X -> X^T X accumulation over a device-resident batch.
//
// Structures are processed in sub-batches whose X rows (bounded to a few GiB)
// are built by the column-block kernel straight into a row-major buffer with
// the row weights applied. The weighted targets are placed in column F of the
// same buffer, and the Gram kernel folds the whole [X | y] block into the
// FP64 accumulator, so X^T y and y^T y come from exactly the X that X^T X
// sees. X for the whole dataset is never materialised. Mirrors the per-batch
// work of _compute_products_single_batch
// (src/pypolymlp/mlp_dev/core/data_sequential.py:97-155).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "capi_internal.hpp"
#include "profile.hpp"

namespace plm {

static constexpr size_t kXBudget = 3ull << 30;   // bytes of X per sub-batch (PLM_X_BUDGET_MB)
static constexpr size_t kABudget = 8ull << 30;   // bytes of per-atom work (a, W', x) per sub-batch
static constexpr size_t kPBudget = 6ull << 30;   // bytes of pair records per sub-batch
static constexpr size_t kDBudget = 4ull << 30;   // bytes of factor derivatives per sub-batch

// Sub-batches run in order on the library stream: X rows (weights applied,
// targets in column F), then the Gram of [X | y]. (Running the Gram of one
// sub-batch on a second stream while the next X is built was measured slower
// for the mixed form: the two contend for the CUs' registers and LDS.)
void normal_accumulate(NormalHandle& acc, BatchHandle& b, const double* row_w, const double* y_w,
                       bool fp64) {
    ModelHandle& m = *acc.model;
    Device& dev = device();
    hipStream_t st = dev.stream;
    const int elem = fp64 ? 8 : 4;
    const int F = acc.F, ld = acc.F_pad;
    const int kstep = 64;  // rows_pad multiple of the Gram's 64-row stage
    const size_t row_bytes = (size_t)ld * elem;
    const size_t rec_bytes = pair_record_bytes(m, fp64);
    const size_t atom_bytes = atom_work_bytes(m, fp64);
    double* xe_sum = reinterpret_cast<double*>(acc.vec.ptr);
    double* xe_sq = xe_sum + acc.F_pad;
    // PLM_X_BUDGET_MB lowers the X budget of a sub-batch (less device memory,
    // more sub-batches)
    size_t x_budget = kXBudget;
    if (const char* e = std::getenv("PLM_X_BUDGET_MB")) {
        const long mb = std::atol(e);
        if (mb > 0) x_budget = std::min(kXBudget, (size_t)mb << 20);
    }
    std::vector<std::vector<int>> rmaps;  // row maps, alive until the stream is synchronised
    int s0 = 0;
    while (s0 < b.n_st) {
        int s1 = s0 + 1;
        auto fits = [&](int s_end) {
            const size_t rows = (size_t)(b.row_off[s_end] - b.row_off[s0]);
            const size_t atoms = (size_t)(b.atom_off[s_end] - b.atom_off[s0]);
            const size_t pairs = (size_t)(b.nb_off[b.atom_off[s_end]] - b.nb_off[b.atom_off[s0]]);
            return (rows + kstep) * row_bytes <= x_budget &&
                   atoms * atom_bytes <= kABudget && pairs * rec_bytes <= kPBudget &&
                   poly_bytes(m, b, s0, s_end, fp64) <= kDBudget;
        };
        while (s1 < b.n_st && fits(s1 + 1)) ++s1;
        const long r0 = b.row_off[s0], r1 = b.row_off[s1];
        const long rows = r1 - r0;
        const long rows_pad = (rows + kstep - 1) / kstep * kstep;
        acc.xbuf.alloc((size_t)rows_pad * row_bytes);
        acc.wbuf.alloc((size_t)rows * 8);
        acc.ybuf.alloc((size_t)rows * 8);
        check_hip(hipMemsetAsync(acc.xbuf.ptr, 0, (size_t)rows_pad * row_bytes, st), "memset");
        check_hip(hipMemcpyAsync(acc.wbuf.ptr, row_w + r0, (size_t)rows * 8, hipMemcpyHostToDevice, st), "copy");
        check_hip(hipMemcpyAsync(acc.ybuf.ptr, y_w + r0, (size_t)rows * 8, hipMemcpyHostToDevice, st), "copy");
        build_x_rows(m, b, s0, s1, fp64, acc.xbuf.ptr, ld, (const double*)acc.wbuf.ptr, xe_sum, xe_sq,
                     acc.work, true);
        launch_set_column(st, fp64, acc.xbuf.ptr, ld, F, (const double*)acc.ybuf.ptr, (int)rows);
        check_hip(hipGetLastError(), "target column");
        if (fp64) {
            ProfScope ps("gram", st);
            launch_gram_f64(st, reinterpret_cast<const double*>(acc.xbuf.ptr), ld, (int)rows_pad,
                            acc.F_pad, (double*)acc.xtx.ptr, acc.F_pad);
        } else {
            // quantisation order: the energy and stress rows (7 per structure),
            // then the force rows; the first exponent chunk is shortened so that
            // the energy and stress rows end on a chunk boundary. An energy row
            // sums a feature over the structure's atoms, so sharing a chunk with
            // it would cost the force rows about log2(N) bits.
            const long chunk = gram_chunk_rows();
            const long nes = 7L * (s1 - s0), nes_pad = (nes + kstep - 1) / kstep * kstep;
            const long h0 = nes_pad % chunk == 0 ? chunk : nes_pad % chunk;
            const long rows_q = nes_pad + (rows - nes + kstep - 1) / kstep * kstep;
            rmaps.emplace_back(rows_q, -1);
            std::vector<int>& rm = rmaps.back();
            long qe = 0, qf = nes_pad;
            for (int s = s0; s < s1; ++s) {
                const long base = b.row_off[s] - r0, n = b.row_off[s + 1] - b.row_off[s];
                for (long k = 0; k < 7; ++k) rm[qe++] = (int)(base + k);
                for (long k = 7; k < n; ++k) rm[qf++] = (int)(base + k);
            }
            acc.rbuf.alloc((size_t)rows_q * sizeof(int));
            check_hip(hipMemcpyAsync(acc.rbuf.ptr, rm.data(), (size_t)rows_q * sizeof(int),
                                     hipMemcpyHostToDevice, st), "copy");
            acc.qbuf.alloc(gram_quant_bytes((int)rows_q, (int)h0, ld));
            {
                ProfScope ps("quantize", st);
                launch_gram_quantize(st, reinterpret_cast<const float*>(acc.xbuf.ptr),
                                     (const int*)acc.rbuf.ptr, ld, (int)rows_q, (int)h0,
                                     acc.F_pad, acc.qbuf.ptr);
            }
            check_hip(hipGetLastError(), "quantize launch");
            ProfScope ps("gram", st);
            launch_gram_i8(st, ld, (int)rows_q, (int)h0, acc.F_pad, acc.qbuf.ptr,
                           (double*)acc.xtx.ptr, acc.F_pad);
        }
        check_hip(hipGetLastError(), "gram launch");
        acc.n_rows += rows;
        s0 = s1;
    }
    check_hip(hipStreamSynchronize(st), "sync");
}

}  // namespace plm
