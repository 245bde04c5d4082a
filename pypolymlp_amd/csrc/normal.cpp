// Fused X -> X^T X accumulation over a device-resident batch.
//
// Structures are processed in sub-batches whose X rows (bounded to a few GiB)
// are built by the column kernels straight into a row-major buffer with the
// row weights applied. The weighted targets are placed in column F of the
// same buffer, and the Gram kernel folds the whole [X | y] block into the
// FP64 accumulator, so X^T y and y^T y come from exactly the X that X^T X
// sees. X for the whole dataset is never materialised. Mirrors the per-batch
// work of _compute_products_single_batch
// (src/pypolymlp/mlp_dev/core/data_sequential.py:97-155).
//
// The work is split into a host-side plan (sub-batch bounds, quantisation row
// maps, factor-derivative offsets, the row weights and targets, all uploaded
// once) and a launch-only sequence over the plan, so that one accumulation
// pass can run eagerly or be captured once as a HIP graph and replayed.
#include <algorithm>
#include <cstdlib>
#include <memory>
#include <cstring>
#include <vector>

#include "capi_internal.hpp"
#include "profile.hpp"

namespace plm {

static constexpr size_t kXBudget = 3ull << 30;   // bytes of X per sub-batch (PLM_X_BUDGET_MB)
static constexpr size_t kABudget = 8ull << 30;   // bytes of per-atom work (a, W', x) per sub-batch
static constexpr size_t kPBudget = 6ull << 30;   // bytes of pair records per sub-batch
static constexpr size_t kDBudget = 4ull << 30;   // bytes of factor derivatives per sub-batch
static constexpr long kStep = 64;                // rows_pad multiple of the Gram's 64-row stage

void plan_accumulate(NormalHandle& acc, BatchHandle& b, const double* row_w, const double* y_w,
                     bool fp64, bool quant, AccumPlan& plan) {
    ModelHandle& m = *acc.model;
    hipStream_t st = device().stream;
    const int elem = fp64 ? 8 : 4;
    const size_t row_bytes = (size_t)acc.F_pad * elem;
    const size_t rec_bytes = pair_record_bytes(m, fp64);
    const size_t atom_bytes = atom_work_bytes(m, fp64);
    // PLM_X_BUDGET_MB lowers the X budget of a sub-batch (less device memory,
    // more sub-batches)
    size_t x_budget = kXBudget;
    if (const char* e = std::getenv("PLM_X_BUDGET_MB")) {
        const long mb = std::atol(e);
        if (mb > 0) x_budget = std::min(kXBudget, (size_t)mb << 20);
    }
    plan.fp64 = fp64;
    plan.quant = quant;
    plan.sb.clear();
    std::vector<int> rmaps;
    std::vector<long> doffs;
    const bool use_poly = m.poly().enabled;
    int s0 = 0;
    while (s0 < b.n_st) {
        int s1 = s0 + 1;
        auto fits = [&](int s_end) {
            const size_t rows = (size_t)(b.row_off[s_end] - b.row_off[s0]);
            const size_t atoms = (size_t)(b.atom_off[s_end] - b.atom_off[s0]);
            const size_t pairs = (size_t)(b.nb_off[b.atom_off[s_end]] - b.nb_off[b.atom_off[s0]]);
            return (rows + kStep) * row_bytes <= x_budget &&
                   atoms * atom_bytes <= kABudget && pairs * rec_bytes <= kPBudget &&
                   poly_bytes(m, b, s0, s_end, fp64) <= kDBudget;
        };
        while (s1 < b.n_st && fits(s1 + 1)) ++s1;
        SubBatch sb{};
        sb.s0 = s0, sb.s1 = s1;
        sb.rows = b.row_off[s1] - b.row_off[s0];
        sb.rows_pad = (sb.rows + kStep - 1) / kStep * kStep;
        sb.doff_off = (long)doffs.size();
        if (use_poly) {
            const std::vector<long> d = poly_d_offsets(m, b, s0, s1);
            doffs.insert(doffs.end(), d.begin(), d.end());
        }
        sb.rmap_off = (long)rmaps.size();
        if (quant) {
            // quantisation order: the energy and stress rows (7 per structure),
            // then the force rows; the first exponent chunk is shortened so that
            // the energy and stress rows end on a chunk boundary. An energy row
            // sums a feature over the structure's atoms, so sharing a chunk with
            // it would cost the force rows about log2(N) bits.
            const long chunk = gram_chunk_rows();
            const long r0 = b.row_off[s0];
            const long nes = 7L * (s1 - s0);
            sb.nes_pad = (nes + kStep - 1) / kStep * kStep;
            sb.h0 = sb.nes_pad % chunk == 0 ? chunk : sb.nes_pad % chunk;
            sb.rows_q = sb.nes_pad + (sb.rows - nes + kStep - 1) / kStep * kStep;
            const size_t base_q = rmaps.size();
            rmaps.resize(base_q + sb.rows_q, -1);
            long qe = 0, qf = sb.nes_pad;
            for (int s = s0; s < s1; ++s) {
                const long base = b.row_off[s] - r0, n = b.row_off[s + 1] - b.row_off[s];
                for (long k = 0; k < 7; ++k) rmaps[base_q + qe++] = (int)(base + k);
                for (long k = 7; k < n; ++k) rmaps[base_q + qf++] = (int)(base + k);
            }
        }
        plan.sb.push_back(sb);
        s0 = s1;
    }
    plan.n_rows = b.row_off[b.n_st];
    auto up = [&](DevBuf& d, const void* h, size_t bytes) {
        d.alloc(std::max<size_t>(bytes, 16));
        if (bytes) check_hip(hipMemcpyAsync(d.ptr, h, bytes, hipMemcpyHostToDevice, st), "copy");
    };
    up(plan.rmaps, rmaps.data(), rmaps.size() * sizeof(int));
    up(plan.doffs, doffs.data(), doffs.size() * sizeof(long));
    up(plan.w, row_w, (size_t)plan.n_rows * 8);
    up(plan.y, y_w, (size_t)plan.n_rows * 8);
    check_hip(hipStreamSynchronize(st), "sync");  // the host tables go out of scope
}

// Sub-batches run in order on one stream: X rows (weights applied, targets
// in column F), then the Gram of [X | y]. (Running the Gram of one sub-batch
// on a second stream while the next X is built was measured slower for the
// mixed form: the two contend for the CUs' registers and LDS.)
void enqueue_accumulate(NormalHandle& acc, BatchHandle& b, const AccumPlan& plan, AccumWork& ws,
                        double* gram, double* vec, hipStream_t st, const KeepRequest* keep) {
    ModelHandle& m = *acc.model;
    const bool fp64 = plan.fp64;
    const int elem = fp64 ? 8 : 4;
    const int F = acc.F, ld = acc.F_pad;
    const size_t row_bytes = (size_t)ld * elem;
    double* xe_sum = vec;
    double* xe_sq = vec + acc.F_pad;
    const double* w = (const double*)plan.w.ptr;
    const double* y = (const double*)plan.y.ptr;
    for (const SubBatch& sb : plan.sb) {
        const long r0 = b.row_off[sb.s0];
        const size_t xbytes = (size_t)sb.rows_pad * row_bytes;
        std::unique_ptr<KeptRows> kr;
        if (keep && acc.kept_bytes + xbytes <= acc.keep_budget) {
            kr = std::make_unique<KeptRows>();
            kr->batch = keep->batch, kr->wy_hash = keep->wy_hash;
            kr->s0 = sb.s0, kr->s1 = sb.s1, kr->fp64 = fp64;
            kr->x.alloc(xbytes);
        } else {
            ws.xbuf.alloc(xbytes);
        }
        void* xb = kr ? kr->x.ptr : ws.xbuf.ptr;
        // cleared by a kernel (a memset node would not replay correctly in a graph)
        launch_fill_zero(st, (double*)xb, xbytes / 8);
        build_x_rows(m, b, sb.s0, sb.s1, fp64, xb, ld, w + r0, xe_sum, xe_sq, ws.work, true,
                     (const long*)plan.doffs.ptr + sb.doff_off);
        launch_set_column(st, fp64, xb, ld, F, y + r0, (int)sb.rows);
        check_hip(hipGetLastError(), "target column");
        if (!plan.quant) {
            ProfScope ps("gram", st);
            launch_gram_f64(st, reinterpret_cast<const double*>(xb), ld, (int)sb.rows_pad,
                            acc.F_pad, gram, acc.F_pad);
        } else {
            ws.qbuf.alloc(gram_quant_bytes((int)sb.rows_q, (int)sb.h0, ld));
            {
                ProfScope ps("quantize", st);
                launch_gram_quantize(st, fp64, xb, (const int*)plan.rmaps.ptr + sb.rmap_off, ld,
                                     (int)sb.rows_q, (int)sb.h0, acc.F_pad, ws.qbuf.ptr);
            }
            check_hip(hipGetLastError(), "quantize launch");
            ProfScope ps("gram", st);
            launch_gram_i8(st, ld, (int)sb.rows_q, (int)sb.h0, acc.F_pad, ws.qbuf.ptr, gram,
                           acc.F_pad);
        }
        check_hip(hipGetLastError(), "gram launch");
        if (kr) {
            acc.kept_bytes += xbytes;
            acc.kept.push_back(std::move(kr));
        }
    }
}

unsigned long long hash_rows(const double* w, const double* y, long n) {
    unsigned long long h = 1469598103934665603ull;
    auto mix = [&](const double* p) {
        const unsigned char* c = reinterpret_cast<const unsigned char*>(p);
        for (size_t k = 0; k < (size_t)n * 8; ++k) h = (h ^ c[k]) * 1099511628211ull;
    };
    mix(w);
    mix(y);
    return h ^ (unsigned long long)n;
}

void normal_accumulate(NormalHandle& acc, BatchHandle& b, const double* row_w, const double* y_w,
                       bool fp64, bool quant) {
    if (!fp64 && !quant) throw std::invalid_argument("the FP64 Gram needs the FP64 X build");
    hipStream_t st = device().stream;
    acc.factors.clear();  // the system changes
    AccumPlan plan;
    plan_accumulate(acc, b, row_w, y_w, fp64, quant, plan);
    KeepRequest kq{b.serial, 0};
    if (acc.keep_budget > acc.kept_bytes) kq.wy_hash = hash_rows(row_w, y_w, plan.n_rows);
    enqueue_accumulate(acc, b, plan, acc.ws, (double*)acc.xtx.ptr, (double*)acc.vec.ptr, st,
                       acc.keep_budget > acc.kept_bytes ? &kq : nullptr);
    check_hip(hipStreamSynchronize(st), "sync");
    acc.n_rows += plan.n_rows;
}

void normal_keep_rows(NormalHandle& acc, size_t max_bytes) {
    check_hip(hipStreamSynchronize(device().stream), "sync");
    acc.keep_budget = max_bytes;
    if (max_bytes == 0) {
        acc.kept.clear();
        acc.kept_bytes = 0;
        acc.factors.clear();
    }
}

NormalGraph::~NormalGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
}

// One eager pass into scratch sums first (every work buffer reaches its size,
// code objects and model tables are loaded, library handles initialised), then
// the same launches captured with the accumulator's own sums as targets.
void normal_graph_capture(NormalGraph& g, const double* row_w, const double* y_w, bool fp64,
                          bool quant) {
    if (!fp64 && !quant) throw std::invalid_argument("the FP64 Gram needs the FP64 X build");
    NormalHandle& acc = *g.acc;
    hipStream_t st = device().stream;
    plan_accumulate(acc, *g.b, row_w, y_w, fp64, quant, g.plan);
    {
        DevBuf gram, vec;
        gram.alloc(acc.xtx.bytes), vec.alloc(acc.vec.bytes);
        enqueue_accumulate(acc, *g.b, g.plan, g.ws, (double*)gram.ptr, (double*)vec.ptr, st);
        check_hip(hipStreamSynchronize(st), "sync");
    }
    ProfPause pause;
    check_hip(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "begin capture");
    try {
        enqueue_accumulate(acc, *g.b, g.plan, g.ws, (double*)acc.xtx.ptr, (double*)acc.vec.ptr, st);
    } catch (...) {
        hipGraph_t tmp = nullptr;
        (void)hipStreamEndCapture(st, &tmp);
        if (tmp) (void)hipGraphDestroy(tmp);
        throw;
    }
    check_hip(hipStreamEndCapture(st, &g.graph), "end capture");
    check_hip(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0), "graph instantiate");
}

void normal_graph_run(NormalGraph& g, int n_steps) {
    if (n_steps < 1) throw std::invalid_argument("steps must be positive");
    hipStream_t st = device().stream;
    g.acc->factors.clear();  // the system changes
    for (int k = 0; k < n_steps; ++k) check_hip(hipGraphLaunch(g.exec, st), "graph launch");
    check_hip(hipStreamSynchronize(st), "sync");
    g.acc->n_rows += (long)n_steps * g.plan.n_rows;
}

// The same sub-batches and X build as the accumulation, then the residual
// kernels in place of the Gram. Used for one step of iterative refinement of
// a solution from the quantised Gram (the residual of the normal equations
// of X itself, X^T y - X^T X c = X^T (y - X c)) and for exact squared errors.
void normal_residual(NormalHandle& acc, BatchHandle& b, const double* row_w, const double* y_w,
                     bool fp64, const double* coeffs, int n_coef, double* xtr, double* sq) {
    if (n_coef < 1 || n_coef > 8) throw std::invalid_argument("residual pass: 1 to 8 coefficient vectors");
    ModelHandle& m = *acc.model;
    hipStream_t st = device().stream;
    const int F = acc.F, ld = acc.F_pad;
    const int elem = fp64 ? 8 : 4;
    AccumPlan plan;
    plan_accumulate(acc, b, row_w, y_w, fp64, false, plan);
    std::vector<double> hc((size_t)n_coef * ld, 0.0);
    for (int k = 0; k < n_coef; ++k)
        std::memcpy(&hc[(size_t)k * ld], coeffs + (size_t)k * F, (size_t)F * 8);
    DevBuf C, R, S, U;
    C.alloc(hc.size() * 8), R.alloc(hc.size() * 8), S.alloc((size_t)n_coef * 8);
    check_hip(hipMemcpyAsync(C.ptr, hc.data(), hc.size() * 8, hipMemcpyHostToDevice, st), "copy");
    check_hip(hipMemsetAsync(R.ptr, 0, hc.size() * 8, st), "memset");
    check_hip(hipMemsetAsync(S.ptr, 0, (size_t)n_coef * 8, st), "memset");
    long u_rows = 1;
    for (const SubBatch& sb : plan.sb) u_rows = std::max(u_rows, sb.rows);
    U.alloc((size_t)u_rows * n_coef * 8);
    AccumWork& ws = acc.ws;
    const double* w = (const double*)plan.w.ptr;
    const double* y = (const double*)plan.y.ptr;
    const unsigned long long wy = acc.kept.empty() ? 0 : hash_rows(row_w, y_w, plan.n_rows);
    for (const SubBatch& sb : plan.sb) {
        if (sb.rows == 0) continue;
        const long r0 = b.row_off[sb.s0];
        // the rows as the accumulation built them, when kept; else built again
        const void* xb = nullptr;
        for (const auto& k : acc.kept)
            if (k->batch == b.serial && k->wy_hash == wy && k->s0 == sb.s0 && k->s1 == sb.s1 &&
                k->fp64 == fp64)
                xb = k->x.ptr;
        if (!xb) {
            ws.xbuf.alloc((size_t)sb.rows_pad * ld * elem);
            launch_fill_zero(st, (double*)ws.xbuf.ptr, (size_t)sb.rows_pad * ld * elem / 8);
            build_x_rows(m, b, sb.s0, sb.s1, fp64, ws.xbuf.ptr, ld, w + r0, nullptr, nullptr,
                         ws.work, true, (const long*)plan.doffs.ptr + sb.doff_off);
            launch_set_column(st, fp64, ws.xbuf.ptr, ld, F, y + r0, (int)sb.rows);
            check_hip(hipGetLastError(), "target column");
            xb = ws.xbuf.ptr;
        }
        ProfScope ps("residual", st);
        launch_residual_pass(st, fp64, xb, ld, (int)sb.rows, F, (const double*)C.ptr, n_coef,
                             (double*)U.ptr, (double*)R.ptr, (double*)S.ptr);
        check_hip(hipGetLastError(), "residual launch");
    }
    std::vector<double> hr(hc.size()), hs(n_coef);
    check_hip(hipMemcpyAsync(hr.data(), R.ptr, hr.size() * 8, hipMemcpyDeviceToHost, st), "copy");
    check_hip(hipMemcpyAsync(hs.data(), S.ptr, hs.size() * 8, hipMemcpyDeviceToHost, st), "copy");
    check_hip(hipStreamSynchronize(st), "sync");
    for (int k = 0; k < n_coef; ++k) {
        for (int f = 0; f < F; ++f) xtr[(size_t)k * F + f] += hr[(size_t)k * ld + f];
        sq[k] += hs[k];
    }
}

}  // namespace plm
