// libmlpcpp support functions on the device: get_fn, get_ylm and the
// synthetic-order-parameter entry points of PolymlpAPI (features and
// prod_sum of a given a_nlm vector), each running the kernels and device
// functions of the X and evaluation paths.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <utility>
#include <vector>

#include "../../include/polymlp_amd.h"
#include "capi_internal.hpp"

namespace plm {

void launch_get_fn(hipStream_t st, bool fp64, int n, const double* r, const double* par, int n_par,
                   double cutoff, double* fn, double* fd);
void launch_get_ylm(hipStream_t st, bool fp64, int n, const double* xyz, int lmax, double* ylm,
                    double* dylm);
void launch_conj_fill(hipStream_t st, bool fp64, int n_nlm, const int* src, const double* sign,
                      const unsigned char* conj, const double* re, const double* im,
                      void* a_full);

namespace {

void upload(DevBuf& b, const void* src, size_t bytes, hipStream_t st) {
    b.alloc(std::max<size_t>(bytes, 16));
    if (bytes) check_hip(hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, st), "copy");
}

void download(void* dst, const DevBuf& b, size_t bytes, hipStream_t st) {
    check_hip(hipMemcpyAsync(dst, b.ptr, bytes, hipMemcpyDeviceToHost, st), "copy");
}

// complex (re, im) of FP32 or FP64 device values into FP64 host pairs
void download_cpx(double* dst, const DevBuf& b, size_t n_cpx, bool fp64, hipStream_t st) {
    if (fp64) {
        download(dst, b, n_cpx * 16, st);
        check_hip(hipStreamSynchronize(st), "sync");
        return;
    }
    std::vector<float> h(2 * n_cpx);
    download(h.data(), b, h.size() * 4, st);
    check_hip(hipStreamSynchronize(st), "sync");
    for (size_t k = 0; k < h.size(); ++k) dst[k] = h[k];
}

// One atom of type t with the synthetic order parameters: the device batch
// view the per-atom kernels read (types only) and the filled a_full.
// The fill follows compute_anlmtp_conjugate's assignment order: for each
// non-conjugate component k, a[ilocal] = v_k, then a[ilocal_conj] = cc conj(v_k);
// a self-conjugate component (m = 0) keeps the second write. (For order
// parameters of real structures a_{nl0} is real and both writes agree.)
struct OneAtom {
    DevBuf type, re, im, a, src, sign, conj;
    BatchDev bd{};
    OneAtom(ModelHandle& m, int t, const double* anlm_re, const double* anlm_im, bool fp64,
            hipStream_t st) {
        if (!m.built) m.build_device();
        if (t < 0 || t >= m.host.n_type()) throw std::invalid_argument("type out of range");
        if (!m.host.input().gtinv)
            throw std::invalid_argument("synthetic order parameters need a gtinv model");
        const auto& T = m.host.type(t);
        const int n_nlm = (int)T.nlm.size(), n_nc = (int)T.noconj.size();
        std::vector<int> hs(n_nlm, 0);
        std::vector<double> hg(n_nlm, 1.0);
        std::vector<unsigned char> hc(n_nlm, 0);
        for (int k = 0; k < n_nc; ++k) {
            const auto& at = T.nlm[T.noconj[k]];
            hs[at.ilocal] = k, hg[at.ilocal] = 1.0, hc[at.ilocal] = 0;
            hs[at.ilocal_conj] = k, hg[at.ilocal_conj] = at.cc, hc[at.ilocal_conj] = 1;
        }
        upload(type, &t, 4, st);
        upload(re, anlm_re, (size_t)n_nc * 8, st);
        upload(im, anlm_im, (size_t)n_nc * 8, st);
        upload(src, hs.data(), hs.size() * 4, st);
        upload(sign, hg.data(), hg.size() * 8, st);
        upload(conj, hc.data(), hc.size(), st);
        a.alloc((size_t)(m.max_nlm + 1) * 2 * (fp64 ? 8 : 4));
        launch_conj_fill(st, fp64, n_nlm, (const int*)src.ptr, (const double*)sign.ptr,
                         (const unsigned char*)conj.ptr, (const double*)re.ptr,
                         (const double*)im.ptr, a.ptr);
        check_hip(hipGetLastError(), "conjugate fill launch");
        check_hip(hipStreamSynchronize(st), "sync");  // the host tables go out of scope
        bd.n_st = 1;
        bd.n_atoms_total = 1;
        bd.types = (const int*)type.ptr;
    }
};

}  // namespace

}  // namespace plm

using namespace plm;

#define PLM_API_TRY(...)                          \
    try {                                         \
        __VA_ARGS__;                              \
        return PLM_OK;                            \
    } catch (const DeviceError& e) {              \
        g_last_error = e.what();                  \
        return PLM_ERR_DEVICE;                    \
    } catch (const std::invalid_argument& e) {    \
        g_last_error = e.what();                  \
        return PLM_ERR_ARG;                       \
    } catch (const std::exception& e) {           \
        g_last_error = e.what();                  \
        return PLM_ERR_MODEL;                     \
    }

extern "C" {

int plm_get_fn(int n, const double* r, const double* params, int n_params, double cutoff,
               int precision, double* fn, double* fn_d) {
    PLM_API_TRY({
        if (n < 0 || n_params < 0 || n_params > kMaxFn) throw std::invalid_argument("bad sizes");
        if (n == 0 || n_params == 0) return PLM_OK;
        hipStream_t st = device().stream;
        DevBuf dr, dp, df, dd;
        upload(dr, r, (size_t)n * 8, st);
        upload(dp, params, (size_t)n_params * 16, st);
        df.alloc((size_t)n * n_params * 8);
        dd.alloc((size_t)n * n_params * 8);
        launch_get_fn(st, precision != PLM_MIXED, n, (const double*)dr.ptr, (const double*)dp.ptr,
                      n_params, cutoff, (double*)df.ptr, (double*)dd.ptr);
        check_hip(hipGetLastError(), "get_fn launch");
        download(fn, df, (size_t)n * n_params * 8, st);
        download(fn_d, dd, (size_t)n * n_params * 8, st);
        check_hip(hipStreamSynchronize(st), "sync");
    });
}

int plm_get_ylm(int n, const double* xyz, int lmax, int precision, double* ylm, double* dylm) {
    PLM_API_TRY({
        const int nh = (lmax + 1) * (lmax + 2) / 2;
        if (n < 0 || lmax < 0 || nh > kMaxLmHalf) throw std::invalid_argument("bad sizes");
        if (n == 0) return PLM_OK;
        hipStream_t st = device().stream;
        DevBuf dx, dy, dd;
        upload(dx, xyz, (size_t)n * 24, st);
        dy.alloc((size_t)n * nh * 16);
        dd.alloc((size_t)n * nh * 48);
        launch_get_ylm(st, precision != PLM_MIXED, n, (const double*)dx.ptr, lmax,
                       (double*)dy.ptr, (double*)dd.ptr);
        check_hip(hipGetLastError(), "get_ylm launch");
        download(ylm, dy, (size_t)n * nh * 16, st);
        download(dylm, dd, (size_t)n * nh * 48, st);
        check_hip(hipStreamSynchronize(st), "sync");
    });
}

int plm_api_features(plm_model* mp, int type, const double* anlm_re, const double* anlm_im,
                     int precision, double* anlm_full, double* features) {
    PLM_API_TRY({
        if (!mp) throw std::invalid_argument("null model");
        auto& m = *reinterpret_cast<ModelHandle*>(mp);
        const bool fp64 = precision != PLM_MIXED;
        hipStream_t st = device().stream;
        OneAtom one(m, type, anlm_re, anlm_im, fp64, st);
        const int elem = fp64 ? 8 : 4;
        const int w_stride = std::max(m.wz_max_type, 1), x_stride = std::max(m.lin_max_type, 1);
        DevBuf w, x;
        w.alloc((size_t)w_stride * 2 * elem);
        x.alloc((size_t)x_stride * elem);
        launch_wprime(st, fp64, m.md, one.bd, one.a.ptr, m.max_nlm, 0, 1, w.ptr, w_stride, x.ptr,
                      x_stride, m.max_nlm, m.w_maxf, m.v_maxf);
        check_hip(hipGetLastError(), "wprime launch");
        const TypeDev& td = m.md.t[type];
        if (anlm_full) download_cpx(anlm_full, one.a, (size_t)td.n_nlm, fp64, st);
        if (features) {
            std::vector<char> h((size_t)td.n_lin * elem);
            download(h.data(), x, h.size(), st);
            check_hip(hipStreamSynchronize(st), "sync");
            for (int f = 0; f < td.n_lin; ++f)
                features[f] = fp64 ? reinterpret_cast<const double*>(h.data())[f]
                                   : reinterpret_cast<const float*>(h.data())[f];
        }
    });
}

int plm_api_sum_of_prod(plm_model* mp, int type, const double* anlm_re, const double* anlm_im,
                        int precision, double* prod_e, double* prod_f) {
    PLM_API_TRY({
        if (!mp) throw std::invalid_argument("null model");
        auto& m = *reinterpret_cast<ModelHandle*>(mp);
        if (!m.pot_uploaded) throw std::invalid_argument("coefficients not set");
        const bool fp64 = precision != PLM_MIXED;
        hipStream_t st = device().stream;
        OneAtom one(m, type, anlm_re, anlm_im, fp64, st);
        const int elem = fp64 ? 8 : 4;
        const int ps = std::max(1, m.max_heads);
        DevBuf pe, pf;
        pe.alloc((size_t)ps * 2 * elem);
        pf.alloc((size_t)ps * 2 * elem);
        launch_prodsum(st, fp64, m.md, m.ed, one.bd, one.a.ptr, m.max_nlm, pe.ptr, pf.ptr, ps,
                       m.max_lin, 0);
        check_hip(hipGetLastError(), "prodsum launch");
        // the sums are stored by head position; returned in the potential's order
        const int nh = m.ed.t[type].n_heads;
        const std::vector<int>& rank = m.head_rank.at(type);
        std::vector<double> tmp(2 * (size_t)nh);
        for (auto [dst, buf] : {std::pair<double*, const DevBuf*>{prod_e, &pe}, {prod_f, &pf}}) {
            if (!dst) continue;
            download_cpx(tmp.data(), *buf, (size_t)nh, fp64, st);
            for (int k = 0; k < nh; ++k)
                dst[2 * k] = tmp[2 * (size_t)rank[k]], dst[2 * k + 1] = tmp[2 * (size_t)rank[k] + 1];
        }
    });
}

}  // extern "C"
