// Polynomial columns of X, second half: the entry products on MFMA and the
// per-row sums of each column's entries.
//
// For one structure and one factor f, the entries e = (f, u) of f give
//   Z[r][e] = sum_i D[i][f][r] M[i][u_e]        (rows r: 6 stress, then 3N forces)
// a product of D_f (RD x N atoms) with the monomials of the entries gathered
// from M (N x m_f). poly_z_kernel computes it in 64-row x 128-entry tiles
// (4 waves of 2 x 4 16x16x4 MFMA blocks, FP32 or FP64), staging 32 atoms of
// both operands per step in LDS; the entries' monomials are gathered from M
// while staging, so no gathered copy of the monomials is written.
// poly_rows_kernel then forms each force / stress row of the polynomial
// columns, X[r][c] = sum over the column's slots of Z[r][e_s(c)]: one
// workgroup per row reads the row of Z once into LDS (in chunks when it is
// larger than 150 KB) and writes the row's columns contiguously. poly_energy_kernel
// writes the energy rows, sum_i prod_s x_{f_s}(i), and their scale sums.
// Products and column order: Model::model_order2/3 (model.cpp:207-267).
#include <hip/hip_runtime.h>

#include <stdexcept>

#include "device_tables.hpp"
#include "kernels.hpp"

namespace plm {

void check_hip(hipError_t e, const char* what);

typedef float f32x4_z __attribute__((__vector_size__(4 * sizeof(float))));
typedef double f64x4_z __attribute__((__vector_size__(4 * sizeof(double))));

// 16x16x4 MFMA: lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16];
// output element v of lane l is row 4 (l / 16) + v (FP32) or l / 16 + 4 v
// (FP64), column l % 16.
template <class T> struct ZMfma;
template <> struct ZMfma<float> {
    using acc = f32x4_z;
    __device__ static acc mma(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};
template <> struct ZMfma<double> {
    using acc = f64x4_z;
    __device__ static acc mma(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};

constexpr int kZRows = 64;    // rows per tile (2 waves x 32)
constexpr int kZEnt = 128;    // entries per tile (2 waves x 64)
constexpr int kZAtoms = 32;   // atoms per LDS stage
constexpr int kZB = kZAtoms * kZEnt / 256;   // B elements staged per thread
constexpr int kZA = kZAtoms * kZRows / 256;  // A elements staged per thread

// grid: (entry tiles, row tiles, structures of the group). tiles[3 t ..]:
// factor, first and end entry of tile t (within the factor's entries). Each
// wave computes 32 rows x 64 entries (2 x 4 MFMA blocks); the next stage of
// both operands is loaded into registers while the current one is consumed
// from LDS.
template <class T>
__global__ void __launch_bounds__(256) poly_z_kernel(PolyDev pd, BatchDev bd, int s_first,
                                                     int s0, int atom_base,
                                                     const T* __restrict__ D,
                                                     const long* __restrict__ d_off,
                                                     const T* __restrict__ M,
                                                     const int* __restrict__ tiles,
                                                     T* __restrict__ Z, long z_stride) {
    using MM = ZMfma<T>;
    __shared__ T As[kZAtoms][kZRows + 1];
    __shared__ T Bs[kZAtoms][kZEnt + 1];
    const int s = s_first + blockIdx.z;
    const int a0 = bd.st_atom_off[s];
    const int nat = bd.st_atom_off[s + 1] - a0;
    const int RD = 6 + 3 * nat;
    const int r0 = blockIdx.y * kZRows;
    if (r0 >= RD) return;
    const int f = tiles[3 * blockIdx.x], e0 = tiles[3 * blockIdx.x + 1],
              e1 = tiles[3 * blockIdx.x + 2];
    const long KR = (long)pd.K * RD;
    const T* Ds = D + d_off[s - s0] + (long)f * RD;  // element (i, r): Ds[i KR + r]
    const T* Ms = M + (size_t)(a0 - atom_base) * pd.U;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = (wave & 1) * 32, we = (wave >> 1) * 64;
    // staging: B, thread -> entry tid % 128, atoms tid / 128 + 2 j;
    // A, thread -> row tid % 64, atoms tid / 64 + 4 j
    const int be = tid % kZEnt;
    const int ub = e0 + be < e1 ? pd.ent_u[e0 + be] : -1;
    const int ar = tid % kZRows;
    const bool arow = r0 + ar < RD;
    T rb[kZB], ra[kZA];
    auto load = [&](int i0) {
#pragma unroll
        for (int j = 0; j < kZB; ++j) {
            const int i = i0 + tid / kZEnt + 2 * j;
            rb[j] = (ub >= 0 && i < nat) ? Ms[(size_t)i * pd.U + ub] : T(0);
        }
#pragma unroll
        for (int j = 0; j < kZA; ++j) {
            const int i = i0 + tid / kZRows + 4 * j;
            ra[j] = (arow && i < nat) ? Ds[(long)i * KR + r0 + ar] : T(0);
        }
    };
    typename MM::acc acc[2][4];
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[m][q] = typename MM::acc{T(0), T(0), T(0), T(0)};
    load(0);
    for (int i0 = 0; i0 < nat; i0 += kZAtoms) {
        __syncthreads();  // the previous stage has been consumed
#pragma unroll
        for (int j = 0; j < kZB; ++j) Bs[tid / kZEnt + 2 * j][be] = rb[j];
#pragma unroll
        for (int j = 0; j < kZA; ++j) As[tid / kZRows + 4 * j][ar] = ra[j];
        __syncthreads();
        if (i0 + kZAtoms < nat) load(i0 + kZAtoms);
#pragma unroll
        for (int k0 = 0; k0 < kZAtoms; k0 += 4) {
            const int k = k0 + (lane >> 4);
            T a[2], b[4];
#pragma unroll
            for (int m = 0; m < 2; ++m) a[m] = As[k][wr + 16 * m + (lane & 15)];
#pragma unroll
            for (int q = 0; q < 4; ++q) b[q] = Bs[k][we + 16 * q + (lane & 15)];
#pragma unroll
            for (int m = 0; m < 2; ++m)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[m][q] = MM::mma(a[m], b[q], acc[m][q]);
        }
    }
    T* Zs = Z + (size_t)blockIdx.z * z_stride;
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = e0 + we + 16 * q + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = r0 + wr + 16 * m + MM::row(lane, v);
                if (r < RD && e < e1) Zs[(size_t)r * pd.E + e] = acc[m][q][v];
            }
        }
}

constexpr int kRowChunkBytes = 150 * 1024;  // LDS per workgroup for one chunk of a row of Z
constexpr int kRowThreads = 1024;

// grid: (rows of the largest structure, structures of the group). Row r of
// structure s: X[row 1 + r][c0 + c] = w sum_s Z[s][r][e_s(c)]. The row of Z
// goes to LDS (coalesced), then each thread sums its columns' entries from
// LDS, four columns' entry records (one 16-byte load each) in flight at once.
template <class T>
__global__ void __launch_bounds__(kRowThreads) poly_rows_kernel(PolyDev pd, BatchDev bd,
                                                                int s_first,
                                                                const T* __restrict__ Z,
                                                                long z_stride, T* __restrict__ X,
                                                                int ld, int row_base,
                                                                const double* __restrict__ roww) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    T* zr = reinterpret_cast<T*>(smem_raw);
    const int s = s_first + blockIdx.y;
    const int nat = bd.st_atom_off[s + 1] - bd.st_atom_off[s];
    const int r = blockIdx.x;
    if (r >= 6 + 3 * nat) return;
    const int row = bd.st_row_off[s] - row_base + 1 + r;
    const T w = T(roww ? roww[row] : 1.0);
    const T* zs = Z + (size_t)blockIdx.y * z_stride + (size_t)r * pd.E;
    T* xr = X + (size_t)row * ld + pd.c0;
    constexpr int CH = kRowChunkBytes / sizeof(T);
    const int tid = threadIdx.x;
    for (int e0 = 0; e0 < pd.E; e0 += CH) {
        const int e1 = min(pd.E, e0 + CH);
        __syncthreads();
        // the row of Z into LDS, eight loads in flight per thread
        for (int e = e0 + tid; e < e1; e += 8 * kRowThreads) {
            T v[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int ee = e + k * kRowThreads;
                v[k] = ee < e1 ? zs[ee] : T(0);
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int ee = e + k * kRowThreads;
                if (ee < e1) zr[ee - e0] = v[k];
            }
        }
        __syncthreads();
        auto part = [&](const int4 p) {
            T v = T(0);
            if (p.x >= e0 && p.x < e1) v += zr[p.x - e0];
            if (p.y >= e0 && p.y < e1) v += zr[p.y - e0];
            if (p.z >= e0 && p.z < e1) v += zr[p.z - e0];
            return w * v;
        };
        int c = tid;
        for (; c + 3 * kRowThreads < pd.n_cols; c += 4 * kRowThreads) {
            int4 p[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) p[u] = pd.col_pos4[c + u * kRowThreads];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const T v = part(p[u]);
                T* dst = xr + c + u * kRowThreads;
                if (e0 == 0) *dst = v;
                else *dst += v;
            }
        }
        for (; c < pd.n_cols; c += kRowThreads) {
            const T v = part(pd.col_pos4[c]);
            if (e0 == 0) xr[c] = v;
            else xr[c] += v;
        }
    }
}

// Energy rows of the polynomial columns: X[row][c0 + c] = w sum_i prod_s x_{f_s}(i),
// the raw sums into the scale sums.
template <class T>
__global__ void __launch_bounds__(256) poly_energy_kernel(PolyDev pd, BatchDev bd, int s_first,
                                                          int atom_base, const T* __restrict__ XF,
                                                          T* __restrict__ X, int ld, int row_base,
                                                          const double* __restrict__ roww,
                                                          double* __restrict__ xe_sum,
                                                          double* __restrict__ xe_sq) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = s_first + blockIdx.y;
    if (c >= pd.n_cols) return;
    const int a0 = bd.st_atom_off[s], nat = bd.st_atom_off[s + 1] - a0;
    const int* ph = pd.col_phi + 3 * (size_t)c;
    const int nf = ph[2] >= 0 ? 3 : 2;
    T e = T(0);
    for (int i = 0; i < nat; ++i) {
        const T* xf = XF + (size_t)(a0 + i - atom_base) * pd.K;
        T p = xf[ph[0]] * xf[ph[1]];
        if (nf == 3) p *= xf[ph[2]];
        e += p;
    }
    const int rb = bd.st_row_off[s] - row_base;
    const double w = roww ? roww[rb] : 1.0;
    X[(size_t)rb * ld + pd.c0 + c] = T(w * double(e));
    if (xe_sum) {
        atomicAdd(xe_sum + pd.c0 + c, double(e));
        atomicAdd(xe_sq + pd.c0 + c, double(e) * double(e));
    }
}

template <class T>
static void poly_z_t(hipStream_t st, const PolyDev& pd, const BatchDev& bd, int s_first, int n_st,
                     int s0, int atom_base, int rd, const void* D, const long* d_off,
                     const void* M, void* Z, long z_stride) {
    const dim3 gz(pd.n_ztiles, (rd + kZRows - 1) / kZRows, n_st);
    poly_z_kernel<T><<<gz, 256, 0, st>>>(pd, bd, s_first, s0, atom_base,
                                         reinterpret_cast<const T*>(D), d_off,
                                         reinterpret_cast<const T*>(M), pd.ztiles,
                                         reinterpret_cast<T*>(Z), z_stride);
}

template <class T>
static void poly_rows_t(hipStream_t st, const PolyDev& pd, const BatchDev& bd, int s_first,
                        int n_st, int atom_base, int rd, const void* XF, const void* Z,
                        long z_stride, void* X, int ld, int row_base, const double* roww,
                        double* xe_sum, double* xe_sq) {
    const size_t sm = std::min<size_t>((size_t)pd.E * sizeof(T), kRowChunkBytes);
    check_hip(hipFuncSetAttribute(reinterpret_cast<const void*>(&poly_rows_kernel<T>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm),
              "polynomial rows LDS attribute");
    poly_rows_kernel<T><<<dim3(rd, n_st), kRowThreads, sm, st>>>(pd, bd, s_first,
                                                         reinterpret_cast<const T*>(Z), z_stride,
                                                         reinterpret_cast<T*>(X), ld, row_base,
                                                         roww);
    poly_energy_kernel<T><<<dim3((pd.n_cols + 255) / 256, n_st), 256, 0, st>>>(
        pd, bd, s_first, atom_base, reinterpret_cast<const T*>(XF), reinterpret_cast<T*>(X), ld,
        row_base, roww, xe_sum, xe_sq);
}

void launch_poly_z(hipStream_t st, bool fp64, const PolyDev& pd, const BatchDev& bd, int s_first,
                   int n_st, int s0, int atom_base, int rd, const void* D, const long* d_off,
                   const void* M, void* Z, long z_stride) {
    if (n_st < 1 || rd < 6 || pd.n_ztiles < 1) throw std::invalid_argument("polynomial entries");
    if (fp64)
        poly_z_t<double>(st, pd, bd, s_first, n_st, s0, atom_base, rd, D, d_off, M, Z, z_stride);
    else
        poly_z_t<float>(st, pd, bd, s_first, n_st, s0, atom_base, rd, D, d_off, M, Z, z_stride);
}

void launch_poly_rows(hipStream_t st, bool fp64, const PolyDev& pd, const BatchDev& bd,
                      int s_first, int n_st, int atom_base, int rd, const void* XF, const void* Z,
                      long z_stride, void* X, int ld, int row_base, const double* roww,
                      double* xe_sum, double* xe_sq) {
    if (n_st < 1 || rd < 6) throw std::invalid_argument("polynomial rows");
    if (fp64)
        poly_rows_t<double>(st, pd, bd, s_first, n_st, atom_base, rd, XF, Z, z_stride, X, ld,
                            row_base, roww, xe_sum, xe_sq);
    else
        poly_rows_t<float>(st, pd, bd, s_first, n_st, atom_base, rd, XF, Z, z_stride, X, ld,
                           row_base, roww, xe_sum, xe_sq);
}

}  // namespace plm
