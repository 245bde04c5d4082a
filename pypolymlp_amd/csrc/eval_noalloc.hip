// Evaluation of one structure with no device allocation: every device buffer of
// the call (inputs staged from the host, both neighbour lists, order
// parameters, per-atom contractions, outputs) is carved out of one file-scope
// device array. For callers that cannot allocate device memory inside their
// loop (serving loops, code under stream capture by its owner, harnesses that
// forbid hipMalloc). The kernels are those of plm_eval; only the buffers differ.
//
// The model must already hold its coefficients (plm_model_set_coeffs, which
// uploads the potential tables once at set-up). Calls on one GPU share the
// array: concurrent calls on the same GPU race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "capi_internal.hpp"
#include "cell.hpp"
#include "profile.hpp"

namespace plm {

constexpr size_t kNoallocScratchBytes = size_t(64) << 20;  // 64 MiB per GPU
constexpr int kNoallocMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {

std::mutex g_scratch_mu;
void* g_scratch_dev[kNoallocMaxDevices] = {};

// Device address of this GPU's copy of `scratch`, looked up once per device
// with that device current. The array's name is only a host-side stand-in.
unsigned char* scratch_base() {
    int dev = 0;
    check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kNoallocMaxDevices) throw DeviceError("device id out of range");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!p) throw DeviceError("hipGetSymbolAddress(scratch) returned null");
        g_scratch_dev[dev] = p;
    }
    return static_cast<unsigned char*>(g_scratch_dev[dev]);
}

// Bump allocator over the scratch array: offsets only, checked against the size.
struct Arena {
    unsigned char* base;
    size_t off = 0;
    void* take(size_t bytes) {
        off = (off + 255) & ~size_t(255);
        if (bytes > kNoallocScratchBytes || off > kNoallocScratchBytes - bytes)
            throw std::invalid_argument(
                "structure too large for plm_eval_noalloc: needs more than " +
                std::to_string(kNoallocScratchBytes >> 20) + " MiB of scratch");
        void* p = base + off;
        off += bytes;
        return p;
    }
};

}  // namespace

void eval_single_noalloc(ModelHandle& m, const double* axis, const double* pos_c,
                         const int* types_in, int n, bool fp64, double* energy, double* forces,
                         double* stress) {
    if (n <= 0) throw std::invalid_argument("structure with no atoms");
    if (!m.built || !m.pot_uploaded)
        throw std::invalid_argument("plm_eval_noalloc needs a model with coefficients set");
    for (int i = 0; i < n; ++i)
        if (types_in[i] < 0 || types_in[i] >= m.host.n_type())
            throw std::invalid_argument("atom type out of range");
    ProfPause no_events;  // no event pairs (they would be created per call)
    hipStream_t st = device().stream;
    const double cutoff = m.host.input().cutoff;
    const double cut2 = cutoff * cutoff, tol = 1e-10, tol2 = tol * tol;
    const int elem = fp64 ? 8 : 4;

    // host side: translations and search positions (NeighborCell)
    std::vector<double> trans, prev;
    cell_translations(axis, pos_c, n, cutoff, trans, prev);
    const int n_tr = (int)(trans.size() / 3);
    std::vector<double> pos((size_t)3 * n);
    for (int j = 0; j < n; ++j)
        for (int a = 0; a < 3; ++a) pos[3 * (size_t)j + a] = prev[(size_t)a * n + j];
    const int tr_off[2] = {0, n_tr}, atom_off[2] = {0, n}, row_off[2] = {0, 7 + 3 * n};
    std::vector<int> zeros(n, 0);

    // fixed part: inputs and the two count arrays
    Arena ar{scratch_base()};
    int* d_types = (int*)ar.take((size_t)n * 4);
    double* d_pos = (double*)ar.take((size_t)n * 24);
    double* d_trans = (double*)ar.take(std::max<size_t>(trans.size(), 3) * 8);
    int* d_tr_off = (int*)ar.take(8);
    int* d_atom_st = (int*)ar.take((size_t)n * 4);
    int* d_atom_off = (int*)ar.take(8);
    int* d_row_off = (int*)ar.take(8);
    int* d_cnt = (int*)ar.take((size_t)n * 8);  // full counts, then half counts
    // the per-atom work buffers are sized before any neighbour is counted, so
    // that an oversized structure is refused before the GPU is touched
    const size_t heads = (size_t)std::max(1, m.max_heads);
    const size_t n_out = 1 + 3 * (size_t)n + 6;
    const size_t fixed_end = ar.off;
    {
        Arena probe = ar;
        probe.take((size_t)n * m.max_nlm * 2 * elem);
        probe.take((size_t)n * heads * 2 * elem);
        probe.take((size_t)n * heads * 2 * elem);
        probe.take(n_out * 8);
        probe.take((size_t)(n + 1) * 8);
    }
    auto up = [&](void* dst, const void* src, size_t bytes) {
        check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st), "copy");
    };
    up(d_types, types_in, (size_t)n * 4);
    up(d_pos, pos.data(), pos.size() * 8);
    if (!trans.empty()) up(d_trans, trans.data(), trans.size() * 8);
    up(d_tr_off, tr_off, 8);
    up(d_atom_st, zeros.data(), (size_t)n * 4);
    up(d_atom_off, atom_off, 8);
    up(d_row_off, row_off, 8);
    int* cnt_f = d_cnt;
    int* cnt_h = d_cnt + n;
    launch_neighbor_full(st, n, d_atom_st, d_atom_off, d_pos, d_tr_off, d_trans, cut2, tol2, cnt_f,
                         nullptr, nullptr, nullptr, nullptr);
    check_hip(hipGetLastError(), "neighbor launch");
    launch_neighbor_half(st, n, d_atom_st, d_atom_off, d_pos, d_tr_off, d_trans, cut2, tol2, tol,
                         cnt_h, nullptr, nullptr, nullptr, nullptr);
    check_hip(hipGetLastError(), "neighbor launch");
    std::vector<int> c((size_t)2 * n);
    check_hip(hipMemcpyAsync(c.data(), d_cnt, (size_t)n * 8, hipMemcpyDeviceToHost, st), "copy");
    check_hip(hipStreamSynchronize(st), "sync");
    std::vector<int> off_f(n + 1, 0), off_h(n + 1, 0);
    for (int i = 0; i < n; ++i) {
        off_f[i + 1] = off_f[i] + c[i];
        off_h[i + 1] = off_h[i] + c[n + i];
    }
    const size_t Mf = std::max(off_f[n], 1), Mh = std::max(off_h[n], 1);

    ar.off = fixed_end;
    int* d_nb_off = (int*)ar.take((size_t)(n + 1) * 4);
    int* d_nb_j = (int*)ar.take(Mf * 4);
    int* d_nb_t = (int*)ar.take(Mf * 4);
    double* d_nb_d = (double*)ar.take(Mf * 24);
    int* d_hb_off = (int*)ar.take((size_t)(n + 1) * 4);
    int* d_hb_j = (int*)ar.take(Mh * 4);
    int* d_hb_t = (int*)ar.take(Mh * 4);
    double* d_hb_d = (double*)ar.take(Mh * 24);
    void* a_buf = ar.take((size_t)n * m.max_nlm * 2 * elem);
    void* pse = ar.take((size_t)n * heads * 2 * elem);
    void* psf = ar.take((size_t)n * heads * 2 * elem);
    double* out = (double*)ar.take(n_out * 8);

    up(d_nb_off, off_f.data(), (size_t)(n + 1) * 4);
    up(d_hb_off, off_h.data(), (size_t)(n + 1) * 4);
    launch_neighbor_full(st, n, d_atom_st, d_atom_off, d_pos, d_tr_off, d_trans, cut2, tol2,
                         nullptr, d_nb_off, d_nb_j, d_nb_t, d_nb_d);
    check_hip(hipGetLastError(), "neighbor launch");
    launch_neighbor_half(st, n, d_atom_st, d_atom_off, d_pos, d_tr_off, d_trans, cut2, tol2, tol,
                         nullptr, d_hb_off, d_hb_j, d_hb_t, d_hb_d);
    check_hip(hipGetLastError(), "neighbor launch");

    BatchDev bd;
    bd.n_st = 1;
    bd.n_atoms_total = n;
    bd.st_atom_off = d_atom_off;
    bd.st_row_off = d_row_off;
    bd.types = d_types;
    bd.atom_st = d_atom_st;
    bd.pos = d_pos;
    bd.nb_off = d_nb_off;
    bd.nb_j = d_nb_j;
    bd.nb_d = d_nb_d;
    enqueue_eval_dev(m, bd, d_hb_off, d_hb_j, d_hb_d, off_h[n], fp64, a_buf, pse, psf, out, n_out,
                     st);
    std::vector<double> h(n_out);
    check_hip(hipMemcpyAsync(h.data(), out, n_out * 8, hipMemcpyDeviceToHost, st), "copy");
    check_hip(hipStreamSynchronize(st), "sync");
    if (energy) *energy = h[0];
    if (forces) std::memcpy(forces, h.data() + 1, (size_t)3 * n * 8);
    if (stress) std::memcpy(stress, h.data() + 1 + 3 * (size_t)n, 48);
}

}  // namespace plm
