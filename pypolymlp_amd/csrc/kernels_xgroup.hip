// Force and stress rows of the linear columns of X by signature groups, on MFMA.
//
// A linear feature f of atom i has the derivative (per pair i -> j, d = r_j - r_i + T)
//   q_f = d x_f(i) / d d = Re sum_key W'_i[f, key] (f'_n(r) u Y_key + f_n(r) dY_key/dd),
// where the keys of W'[f, .] for the pair's type pair are the same for all the
// features of one signature group. For one atom, one task (16 features of a
// group: the slots) and one type pair, the sums
//   S_c[pair, slot] = sum_key Re Y_c[pair, key] Re W'[key, slot] - Im Y_c Im W',
// c = 0 (Y) and 1..3 (dY/dx, dy, dz), are one matrix product: rows (pair, c),
// inner dimension (key, re/im), columns slot. A wave computes it with 16x16x4
// MFMAs for 4 pairs (16 rows) per tile and 2 keys per step; W' (the B operand)
// is loaded once per atom and type pair. Each lane then holds S_0..S_3 of one
// (pair, slot), so q_c = f'_n u_c S_0 + f_n S_c needs no data exchange; -q goes
// to the rows of atom j (LDS atomics into the structure's row tile), q to the
// rows of atom i and -q d to the stress rows (register sums per lane, added
// with LDS atomics per atom and per structure).
//
// The workgroup holds the stress and force rows of one structure for TB tasks
// in LDS, in FP64 for both precisions (LDS FP64 atomic adds measured several
// times faster than FP32 ones on gfx950 for this scatter); A waves per task
// split the structure's atoms of the task's type.
// One launch per atom type (the tasks of one type own distinct columns), each
// adding its rows into X, which is zero before the first.
//
// Reference loop nests:
//   Local::compute_anlmtp_d            src/cxx/src/compute/local.cpp:116-217
//   Features::compute_features_deriv   src/cxx/src/polymlp/polymlp_features.cpp:208-248
//   PyModel row layout (energy, 6 stress, 3N forces)  src/cxx/src/compute/py_model.cpp:58-106
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <stdexcept>
#include <tuple>

#include "device_tables.hpp"
#include "kernels.hpp"
#include "pair_data.hpp"

namespace plm {

void check_hip(hipError_t e, const char* what);

typedef float f32x4 __attribute__((__vector_size__(4 * sizeof(float))));
typedef double f64x4 __attribute__((__vector_size__(4 * sizeof(double))));

// 16x16x4 MFMA per precision. Lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16]. Output: FP32 row 4 (l / 16) + v, FP64 row l / 16 + 4 v
// (probed on gfx950), so the A row of (pair p, component c) is 4p + c in FP32
// and p + 4c in FP64: either way acc[v] of lane l is (pair l / 16, c = v).
template <class T> struct GroupMfma;
template <> struct GroupMfma<float> {
    using acc = f32x4;
    static constexpr int waves = 16;  // waves per workgroup (register budget)
    __device__ static acc mma(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static int a_pair(int i) { return i >> 2; }
    __device__ static int a_comp(int i) { return i & 3; }
};
template <> struct GroupMfma<double> {
    using acc = f64x4;
    static constexpr int waves = 8;
    __device__ static acc mma(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static int a_pair(int i) { return i & 3; }
    __device__ static int a_comp(int i) { return i >> 2; }
};

constexpr int kGroupKSteps = 6;  // key steps (2 keys each) held per atom and type pair

// four consecutive record elements in one load (records are 4-element aligned)
template <class T> struct GroupVec4;
template <> struct GroupVec4<float> { using type = float4; };
template <> struct GroupVec4<double> { using type = double4; };

// Per (atom, type pair) state of a wave: the B operand of each key step and
// the A-operand offset of this lane in a pair record (-1: the constant Y_00,
// -2: padding).
template <class T>
struct GroupKeys {
    T b[kGroupKSteps];
    int off[kGroupKSteps];
};

// A operand of this lane for key step s: component of Y of its pair
template <class T>
__device__ __forceinline__ T a_value(const T* rec, int off, bool const_one) {
    const T v = rec[max(off, 0)];
    return off >= 0 ? v : ((off == -1 && const_one) ? T(1) : T(0));
}

// One group of 16 pairs [pc, pc + 16) of the segment ending at pe, KS key steps.
template <class T, int KS>
__device__ __forceinline__ void group_pairs(const T* __restrict__ pd, long pair_base, int stride,
                                            int nfn, int pc, int pe, const GroupKeys<T>& K,
                                            int lane, int nid, double* tile, int ldt, int col,
                                            bool keep, T (&self)[3], T (&st)[6]) {
    using M = GroupMfma<T>;
    const int i = lane & 15;
    const bool const_one = M::a_comp(i) == 0 && ((lane >> 4) & 1) == 0;
    // A operands of the 4 tiles (4 pairs each) and the combine's per-pair values
    T a[4][KS];
    T dx[4], dy[4], dz[4], r[4], valid[4], jl[4], fv[4], fd[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int pa = min(pc + 4 * m + M::a_pair(i), pe - 1);
        const T* ra = pd + (size_t)(pa - pair_base) * stride;
#pragma unroll
        for (int s = 0; s < KS; ++s) a[m][s] = a_value<T>(ra, K.off[s], const_one);
        const int pb = min(pc + 4 * m + (lane >> 4), pe - 1);
        const T* rb = pd + (size_t)(pb - pair_base) * stride;
        using V4 = typename GroupVec4<T>::type;
        const V4 h0 = *reinterpret_cast<const V4*>(rb);      // d, r
        const V4 h1 = *reinterpret_cast<const V4*>(rb + 4);  // valid, tp, j
        dx[m] = h0.x, dy[m] = h0.y, dz[m] = h0.z, r[m] = h0.w, valid[m] = h1.x, jl[m] = h1.z;
        const int n = max(nid, 0);
        fv[m] = rb[8 + n];
        fd[m] = rb[8 + nfn + n];
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        if (pc + 4 * m >= pe) break;
        typename M::acc acc = {T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int s = 0; s < KS; ++s) acc = M::mma(a[m][s], K.b[s], acc);
        // q_c = f'_n u_c S_0 + f_n S_c; pairs beyond the cutoff and f_n < 1e-20
        // give nothing (local.cpp:164)
        const bool ok = pc + 4 * m + (lane >> 4) < pe && valid[m] != T(0) && nid >= 0 &&
                        fv[m] >= T(1e-20);
        const T g = ok ? fd[m] * acc[0] / r[m] : T(0);
        const T f = ok ? fv[m] : T(0);
        const T qx = g * dx[m] + f * acc[1];
        const T qy = g * dy[m] + f * acc[2];
        const T qz = g * dz[m] + f * acc[3];
        if (ok && keep) {
            double* rj = tile + (size_t)(6 + 3 * int(jl[m])) * ldt + col;
            atomicAdd(rj, -double(qx));
            atomicAdd(rj + ldt, -double(qy));
            atomicAdd(rj + 2 * ldt, -double(qz));
        }
        self[0] += qx, self[1] += qy, self[2] += qz;
        st[0] -= qx * dx[m], st[1] -= qy * dy[m], st[2] -= qz * dz[m];
        st[3] -= qx * dy[m], st[4] -= qy * dz[m], st[5] -= qz * dx[m];
    }
}

// One item: structure s, task block blk (TB tasks), slot part `part` (of
// P.split). The tile holds the item's stress and force rows for its TC
// columns; the slot lanes outside the part compute and drop their sums.
template <class T>
__device__ __forceinline__ void xgroup_item(const GroupDev& gd, const BatchDev& bd,
                                            const XGroupParams& P, const T* __restrict__ wp,
                                            const T* __restrict__ pd, const PairRec& R,
                                            const int* __restrict__ tpo, T* __restrict__ X,
                                            const double* __restrict__ roww, double* tile, int s,
                                            int blk, int part) {
    const int a0 = bd.st_atom_off[s], nat = bd.st_atom_off[s + 1] - a0;
    const int rows = 6 + 3 * nat;  // stress rows, then 3 force rows per atom
    const int SPP = kGroupSlots / P.split;  // slots per part
    const int TC = P.TB * SPP, ldt = TC + 1;
    const int tid = threadIdx.x;
    for (int e = tid; e < rows * ldt; e += blockDim.x) tile[e] = 0.0;
    __syncthreads();

    // the wave index is uniform in the wave: readfirstlane lets the compiler
    // keep everything derived from it (task, atom, key counts) in scalar registers
    const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int tw = wave % P.TB, stream = wave / P.TB;
    const int task = P.vt0 + blk * P.TB + tw;
    const int ntp = gd.ntp;
    const int j = lane & 15;                  // slot of this lane in the output
    const int kq = lane >> 4;                 // k index of this lane in the A / B operands
    const int kh = kq >> 1, reim = kq & 1;    // key of the step, real / imaginary part
    const int ac = GroupMfma<T>::a_comp(lane & 15);
    const bool keep = j / SPP == part;        // this lane's slot is in the item's part
    if (task < P.vt1 && stream < P.A) {
        const int t = gd.task_type[task];
        const int col = tw * SPP + j - part * SPP;
        // per type pair tables of the task, one type pair per lane: read with
        // readlane, so the atom loop waits on no scalar load (scalar loads
        // share their counter with the LDS atomics)
        const int tl = task * ntp + min(lane, ntp - 1);
        const int v_nkey = gd.tt_nkey[tl], v_woff = gd.tt_woff[tl], v_koff = gd.tt_keyoff[tl];
        T st[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
        int nth = 0;
        for (int base = 0; base < nat; base += 64) {
            // atoms of the task's type among the next 64, in order
            const int ty = base + lane < nat ? bd.types[a0 + base + lane] : -1;
            unsigned long long mask = __ballot(ty == t);
            while (mask) {
                const int ia = base + __builtin_ctzll(mask);
                mask &= mask - 1;
                if (nth++ % P.A != stream) continue;
                const int i = a0 + ia;
                T self[3] = {T(0), T(0), T(0)};
                const T* wi = wp + (size_t)(i - P.atom_base) * P.w_stride * 2;  // (re, im)
                const int v_to = tpo[(size_t)(i - P.atom_base) * (ntp + 1) + min(lane, ntp)];
                for (int tp = 0; tp < ntp; ++tp) {
                    const int nkey = __builtin_amdgcn_readlane(v_nkey, tp);
                    const int pb = __builtin_amdgcn_readlane(v_to, tp);
                    const int pe = __builtin_amdgcn_readlane(v_to, tp + 1);
                    if (nkey == 0 || pb >= pe) continue;
                    const int* keys = gd.keys + __builtin_amdgcn_readlane(v_koff, tp);
                    const int* src = gd.wg_src + (size_t)t * gd.wg_stride +
                                     __builtin_amdgcn_readlane(v_woff, tp);
                    const int nid = gd.tt_nid[(size_t)(task * ntp + tp) * kGroupSlots + j];
                    const int ks = (nkey + 1) >> 1;
                    for (int s0 = 0; s0 < ks; s0 += kGroupKSteps) {
                        // B operand: lane (k = kq, column j) holds Re W' or -Im W'
                        // of key 2 s + kh and slot j; A offsets: this lane's component
                        GroupKeys<T> K;
#pragma unroll
                        for (int q = 0; q < kGroupKSteps; ++q) {
                            const int key_i = 2 * (s0 + q) + kh;
                            const bool real = key_i < nkey;
                            const int kc = real ? key_i : 0;
                            // the W' nonzero at grouped position (key, slot), if any
                            const int g = real ? src[kc * kGroupSlots + j] : -1;
                            const T w = wi[(size_t)max(g, 0) * 2 + reim];
                            K.b[q] = g >= 0 ? (reim ? -w : w) : T(0);
                            const int key = real ? keys[kc] : -2;
                            K.off[q] = key >= 0 ? R.ylm(key) + 2 * ac + reim : key;
                        }
                        const int nks = min(kGroupKSteps, ks - s0);
                        for (int pc = pb; pc < pe; pc += 16) {
                            switch (nks) {
#define PLM_KS(N)                                                                               \
    case N:                                                                                     \
        group_pairs<T, N>(pd, P.pair_base, R.stride, R.nfn, pc, pe, K, lane, nid, tile, ldt, \
                          col, keep, self, st);                                                 \
        break;
                                PLM_KS(1) PLM_KS(2) PLM_KS(3) PLM_KS(4) PLM_KS(5) PLM_KS(6)
#undef PLM_KS
                                default: break;
                            }
                        }
                    }
                }
                // rows of atom i: the four lanes of a slot add their sums
                if (keep) {
                    double* ri = tile + (size_t)(6 + 3 * ia) * ldt + col;
                    atomicAdd(ri, double(self[0]));
                    atomicAdd(ri + ldt, double(self[1]));
                    atomicAdd(ri + 2 * ldt, double(self[2]));
                }
            }
        }
        // stress rows (xx, yy, zz, xy, yz, zx)
        if (keep) {
#pragma unroll
            for (int c = 0; c < 6; ++c) atomicAdd(tile + (size_t)c * ldt + col, double(st[c]));
        }
    }
    __syncthreads();
    // add into X (weights applied per row); tile row r is X row 1 + r
    const int rbase = bd.st_row_off[s] - P.row_base;
    for (int e = tid; e < rows * TC; e += blockDim.x) {
        const int r = e / TC, cc = e % TC;
        const int vv = P.vt0 + blk * P.TB + cc / SPP;
        if (vv >= P.vt1) continue;
        const int g = gd.task_gcol[vv * kGroupSlots + part * SPP + cc % SPP];
        if (g < 0) continue;
        const double w = roww ? roww[rbase + 1 + r] : 1.0;
        T* dst = X + (size_t)(rbase + 1 + r) * P.ld + g;
        *dst += T(w * tile[(size_t)r * ldt + cc]);
    }
    __syncthreads();  // the tile is reused by the workgroup's next item
}

// Static grid: one workgroup per (structure, task block, part).
template <class T>
__global__ void __launch_bounds__(64 * GroupMfma<T>::waves) xgroup_kernel(
    GroupDev gd, BatchDev bd, XGroupParams P, const T* __restrict__ wp, const T* __restrict__ pd,
    PairRec R, const int* __restrict__ tpo, T* __restrict__ X, const double* __restrict__ roww) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    double* tile = reinterpret_cast<double*>(smem_raw);
    const int nbp = P.n_blocks * P.split;  // items per structure
    int item, s;
    if (P.xcd_map) {
        // the blocks of one structure run on one XCD (workgroups are dealt to
        // the 8 XCDs round robin), so its pair records are shared in one L2
        const int xcd = blockIdx.x & 7, k = blockIdx.x >> 3;
        s = P.st0 + (k / nbp) * 8 + xcd;
        item = k % nbp;
        if (s >= P.st_end) return;
    } else {
        item = blockIdx.x % nbp;
        s = P.st0 + blockIdx.x / nbp;
    }
    xgroup_item<T>(gd, bd, P, wp, pd, R, tpo, X, roww, tile, s, item / P.split, item % P.split);
}

// Persistent grid with a work queue: the structures of the launch are dealt
// to 8 queues (structure k to queue k mod 8), each queue hands out its items
// (task block, part) largest structure first (the structures of a batch are
// in ascending size order), and a workgroup drains the queue of its own XCD
// first, so that one structure's items mostly run on one XCD and share its
// pair records in that L2, then takes items from the other queues until all
// are empty. Uneven structure sizes balance over the resident workgroups;
// every workgroup exits when the eight queues are drained.
template <class T>
__global__ void __launch_bounds__(64 * GroupMfma<T>::waves) xgroup_queue_kernel(
    GroupDev gd, BatchDev bd, XGroupParams P, const T* __restrict__ wp, const T* __restrict__ pd,
    PairRec R, const int* __restrict__ tpo, T* __restrict__ X, const double* __restrict__ roww) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    __shared__ int next;
    double* tile = reinterpret_cast<double*>(smem_raw);
    const int nbp = P.n_blocks * P.split;
    const int n_st = P.st_end - P.st0;
    const int xcd = blockIdx.x & 7;
    for (int qq = 0; qq < 8; ++qq) {
        const int q = (xcd + qq) & 7;
        const int nsq = q < n_st ? (n_st - q + 7) / 8 : 0;  // structures of queue q
        const int n_items = nsq * nbp;
        for (;;) {
            if (threadIdx.x == 0) next = atomicAdd(P.queue + q, 1);
            __syncthreads();
            // uniform: kept in scalar registers
            const int k = __builtin_amdgcn_readfirstlane(next);
            __syncthreads();  // read by every thread before thread 0 writes again
            if (k >= n_items) break;
            const int m = nsq - 1 - k / nbp;  // largest structure of the queue first
            const int item = k % nbp;
            xgroup_item<T>(gd, bd, P, wp, pd, R, tpo, X, roww, tile, P.st0 + q + 8 * m,
                           item / P.split, item % P.split);
        }
    }
}

// Energy rows of the linear columns: sum over the structure's atoms of x_f(i)
// in atom order (the feature's local index for the atom's type), weighted; the
// raw values feed the scale sums.
template <class T>
__global__ void __launch_bounds__(256) xenergy_kernel(GroupDev gd, BatchDev bd, int n_lin, int st0,
                                                      int row_base, int atom_base,
                                                      const T* __restrict__ x, int x_stride,
                                                      T* __restrict__ X, int ld,
                                                      const double* __restrict__ roww,
                                                      double* __restrict__ xe_sum,
                                                      double* __restrict__ xe_sq) {
    const int s = st0 + blockIdx.y;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_lin) return;
    T e = T(0);
    for (int i = bd.st_atom_off[s]; i < bd.st_atom_off[s + 1]; ++i) {
        const int l = gd.g2l[(size_t)bd.types[i] * n_lin + c];
        if (l >= 0) e += x[(size_t)(i - atom_base) * x_stride + l];
    }
    const int r = bd.st_row_off[s] - row_base;
    const double w = roww ? roww[r] : 1.0;
    X[(size_t)r * ld + c] = T(w * double(e));
    if (xe_sum) {
        atomicAdd(xe_sum + c, double(e));
        atomicAdd(xe_sq + c, double(e) * double(e));
    }
}

int xgroup_slots(bool) { return kGroupSlots; }

int xgroup_max_waves(bool fp64) {
    return fp64 ? GroupMfma<double>::waves : GroupMfma<float>::waves;
}

size_t xgroup_smem(const XGroupParams& P, int rows_max, bool fp64) {
    const int rows = rows_max - 1;  // the energy row is not in the tile
    return (size_t)rows * (P.TB * xgroup_slots(fp64) / P.split + 1) * sizeof(double);
}

template <class T>
static const void* xgroup_fn(bool persistent) {
    return persistent ? reinterpret_cast<const void*>(&xgroup_queue_kernel<T>)
                      : reinterpret_cast<const void*>(&xgroup_kernel<T>);
}

int xgroup_persistent_grid(const XGroupParams& P, int rows_max, bool fp64) {
    static std::mutex mu;
    static std::map<std::tuple<int, int, size_t>, int> cache;
    const size_t sm = xgroup_smem(P, rows_max, fp64);
    int dev = 0;
    check_hip(hipGetDevice(&dev), "device");
    const auto key = std::make_tuple(dev * 2 + (fp64 ? 1 : 0), 64 * P.TB * P.A, sm);
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int n_cu = 0;
    check_hip(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev),
              "device attribute");
    const void* fn = fp64 ? xgroup_fn<double>(true) : xgroup_fn<float>(true);
    if (sm > 160 * 1024 - 1024) throw std::invalid_argument("xgroup row tile exceeds LDS");
    check_hip(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm),
              "xgroup LDS attribute");
    int per_cu = 0;
    check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64 * P.TB * P.A, sm),
              "occupancy");
    const int g = std::max(1, per_cu) * std::max(1, n_cu);
    cache[key] = g;
    return g;
}

template <class T>
static void launch_xgroup_t(hipStream_t st, const GroupDev& gd, const BatchDev& bd,
                            const XGroupParams& P, int rows_max, const void* wp, const void* pd,
                            const PairRec& R, const int* tpo, void* X, const double* roww) {
    const size_t sm = xgroup_smem(P, rows_max, sizeof(T) == 8);
    const bool persistent = P.queue != nullptr;
    check_hip(hipFuncSetAttribute(xgroup_fn<T>(persistent),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm),
              "xgroup LDS attribute");
    const int n_st = P.st_end - P.st0;
    const long nbp = (long)P.n_blocks * P.split;
    if (persistent) {
        xgroup_queue_kernel<T><<<P.grid, 64 * P.TB * P.A, sm, st>>>(
            gd, bd, P, reinterpret_cast<const T*>(wp), reinterpret_cast<const T*>(pd), R, tpo,
            reinterpret_cast<T*>(X), roww);
        return;
    }
    const long grid = P.xcd_map ? (long)((n_st + 7) / 8) * 8 * nbp : (long)n_st * nbp;
    xgroup_kernel<T><<<grid, 64 * P.TB * P.A, sm, st>>>(
        gd, bd, P, reinterpret_cast<const T*>(wp), reinterpret_cast<const T*>(pd), R, tpo,
        reinterpret_cast<T*>(X), roww);
}

void launch_xgroup(hipStream_t st, bool fp64, const GroupDev& gd, const BatchDev& bd,
                   const XGroupParams& P, int rows_max, const void* wp, const void* pd,
                   const PairRec& R, const int* tpo, void* X, const double* roww) {
    if (P.TB * P.A > xgroup_max_waves(fp64) || P.TB < 1 || P.A < 1)
        throw std::invalid_argument("xgroup launch shape");
    if (!(P.split == 1 || ((P.split == 2 || P.split == 4) && P.TB == 1)))
        throw std::invalid_argument("xgroup slot split");
    if (P.queue && P.grid < 1) throw std::invalid_argument("xgroup persistent grid");
    if (xgroup_smem(P, rows_max, fp64) > 160 * 1024 - 1024)
        throw std::invalid_argument("xgroup row tile exceeds LDS");
    if (fp64)
        launch_xgroup_t<double>(st, gd, bd, P, rows_max, wp, pd, R, tpo, X, roww);
    else
        launch_xgroup_t<float>(st, gd, bd, P, rows_max, wp, pd, R, tpo, X, roww);
}

void launch_xenergy(hipStream_t st, bool fp64, const GroupDev& gd, const BatchDev& bd, int n_lin,
                    int st0, int n_st, int row_base, int atom_base, const void* x, int x_stride,
                    void* X, int ld, const double* roww, double* xe_sum, double* xe_sq) {
    const dim3 grid((n_lin + 255) / 256, n_st);
    if (fp64)
        xenergy_kernel<double><<<grid, 256, 0, st>>>(gd, bd, n_lin, st0, row_base, atom_base,
                                                     reinterpret_cast<const double*>(x), x_stride,
                                                     reinterpret_cast<double*>(X), ld, roww,
                                                     xe_sum, xe_sq);
    else
        xenergy_kernel<float><<<grid, 256, 0, st>>>(gd, bd, n_lin, st0, row_base, atom_base,
                                                    reinterpret_cast<const float*>(x), x_stride,
                                                    reinterpret_cast<float*>(X), ld, roww, xe_sum,
                                                    xe_sq);
}

}  // namespace plm
