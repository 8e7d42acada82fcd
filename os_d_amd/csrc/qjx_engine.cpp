// qjx_engine.cpp -- host side of the MI355X trajectory engine behind include/qjx.h.
//
// Owns the device memory of one dimer model (propagator tree) and one ensemble,
// and drives the grouped step iterations until every trajectory has finished
// the requested periods.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/qjx.h"
#include "qjx_internal.h"
#include "qjx_kernels.h"
#include "qjx_rk.h"
#include "qjx_runner.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(e_ == hipErrorOutOfMemory ? QJX_ERR_NOMEM : QJX_ERR_HIP,         \
                        std::string(#expr) + ": " + hipGetErrorString(e_));              \
    } while (0)

const double kPi = 3.1415926535897932384626433832795;   // reference config.h PI
const uint64_t kMcgA = 302875106592253ULL;              // 13^13

uint64_t mcg_pow_host(uint64_t a, uint64_t e)
{
    const uint64_t mask = (1ULL << 59) - 1ULL;
    uint64_t r = 1;
    while (e) {
        if (e & 1ULL) r = (r * a) & mask;
        a = (a * a) & mask;
        e >>= 1;
    }
    return r;
}

template <typename T>
void dev_free(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

}  // namespace

#include "qjx_ctx.h"

// ---------------------------------------------------------------------------
// Model (newdel.cpp:3-37, 450-498, 1454-1498, 2285-2388)
// ---------------------------------------------------------------------------
static void build_model(qjx_ctx *c)
{
    const qjx_model_desc &m = c->model;
    const int d = c->d;
    const double U = m.prm_U / double(d - 1);
    c->Hdiag.assign(d, 0.0);
    c->Hdrv.assign(d, 0.0);
    c->Ldiag.assign(d, 0.0);
    c->Hoff.assign(std::max(d - 1, 1), 0.0);
    c->Loff.assign(std::max(d - 1, 1), 0.0);
    for (int k = 0; k < d; k++) {
        c->Hdiag[k] = 2.0 * U * double(k * (k - 1) + (d - (k + 1)) * (d - (k + 1) - 1));
        c->Hdrv[k] = double((d - (k + 1)) - k);
        c->Ldiag[k] = double(k - (d - (k + 1)));
    }
    for (int k = 0; k + 1 < d; k++) {
        c->Hoff[k] = 0.0 - m.prm_J * std::sqrt(double((d - (k + 1)) * (k + 1)));
        c->Loff[k] = std::sqrt(double((d - (k + 1)) * (k + 1)));
    }
    // Observables use H + E * H_drv with the static E (experiment.cpp:1446).
    c->HdiagE.assign(d, 0.0);
    for (int k = 0; k < d; k++) c->HdiagE[k] = c->Hdiag[k] + m.prm_E * c->Hdrv[k];

    // Dense L (row-major): L_kk, L_{k+1,k} = -Loff[k], L_{k,k+1} = +Loff[k].
    auto Lel = [&](int i, int j) -> double {
        if (i == j) return c->Ldiag[i];
        if (i == j + 1) return -c->Loff[j];
        if (j == i + 1) return c->Loff[i];
        return 0.0;
    };
    const double g = m.diss_gamma / double(d - 1);
    const double E[2] = {m.prm_E + m.drv_ampl, m.prm_E - m.drv_ampl};
    for (int b = 0; b < 2; b++) c->Hqj[b].assign((size_t)2 * d * d, 0.0);
    for (int i = 0; i < d; i++) {
        for (int j = std::max(0, i - 2); j <= std::min(d - 1, i + 2); j++) {
            // (L^dag L)_ij = sum_k L_ki L_kj, ascending k as the reference zgemm
            double s = 0.0;
            for (int k = std::max(0, std::max(i, j) - 1); k <= std::min(d - 1, std::min(i, j) + 1); k++)
                s += Lel(k, i) * Lel(k, j);
            const double ldl = g * s;
            double h = 0.0, hdrv = 0.0;
            if (i == j) {
                h = c->Hdiag[i];
                hdrv = c->Hdrv[i];
            } else if (i == j + 1) {
                h = c->Hoff[j];
            } else if (j == i + 1) {
                h = c->Hoff[i];
            }
            const double hre = h + 0.5 * 0.0;          // Im(diss_part) = 0
            const double him = 0.0 - 0.5 * ldl;
            for (int b = 0; b < 2; b++) {
                const size_t o = 2 * ((size_t)i * d + j);
                c->Hqj[b][o] = him + E[b] * 0.0;
                c->Hqj[b][o + 1] = 0.0 - (hre + E[b] * hdrv);
            }
        }
    }
}

// Squarings s of the scaling step for exp(A dt): the smallest s >= 0 with
// ||A dt / 2^s||_1 <= 1.
static int expm_squarings(double anorm1, double dt)
{
    const double a1 = anorm1 * dt;
    return a1 > 1.0 ? (int)std::ceil(std::log2(a1)) : 0;
}

// exp(A dt) for the dense complex A on the device by scaling and squaring of a
// degree-18 Taylor polynomial evaluated with the Paterson-Stockmeyer scheme
// (M. S. Paterson, L. J. Stockmeyer, SIAM J. Comput. 2 (1973) 60-66):
// ||A dt / 2^s||_1 <= 1 bounds the truncation error by 1/19! * e < 3e-17.
static int device_expm(qjx_ctx *c, const double *dA, double anorm1, double dt, double *out,
                       double *work)
{
    const int n = c->d;
    const size_t nn2 = (size_t)2 * n * n;
    double *X = work, *X2 = work + nn2, *X3 = work + 2 * nn2, *X4 = work + 3 * nn2,
           *P = work + 4 * nn2, *Tm = work + 5 * nn2;
    const int s = expm_squarings(anorm1, dt);
    hipStream_t st = c->stream;
    qjx::launch_zaxpby_eye(n, dA, std::ldexp(dt, -s), 0.0, X, st);
    qjx::launch_zgemm(n, X, X, X2, st);
    qjx::launch_zgemm(n, X2, X, X3, st);
    qjx::launch_zgemm(n, X2, X2, X4, st);
    double cf[19];
    cf[0] = 1.0;
    for (int j = 1; j <= 18; j++) cf[j] = cf[j - 1] / double(j);
    // P = c16 I + c17 X + c18 X2
    qjx::launch_zaxpby_eye(n, X, cf[17], cf[16], P, st);
    qjx::launch_zaxpy(n, X2, cf[18], P, st);
    for (int blk = 3; blk >= 0; blk--) {
        qjx::launch_zgemm(n, X4, P, Tm, st);
        std::swap(P, Tm);
        // P += c_{4blk} I + c_{4blk+1} X + c_{4blk+2} X2 + c_{4blk+3} X3
        qjx::launch_zaxpby_eye(n, X, cf[4 * blk + 1], cf[4 * blk], Tm, st);
        qjx::launch_zaxpy(n, Tm, 1.0, P, st);
        qjx::launch_zaxpy(n, X2, cf[4 * blk + 2], P, st);
        qjx::launch_zaxpy(n, X3, cf[4 * blk + 3], P, st);
    }
    for (int k = 0; k < s; k++) {
        qjx::launch_zgemm(n, P, P, Tm, st);
        std::swap(P, Tm);
    }
    HIP_TRY(hipMemcpyAsync(out, P, nn2 * sizeof(double), hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipGetLastError());
    return QJX_OK;
}

static double norm1_host(const std::vector<double> &A, int n)
{
    double best = 0.0;
    for (int j = 0; j < n; j++) {
        double s = 0.0;
        for (int i = 0; i < n; i++) {
            const size_t o = 2 * ((size_t)i * n + j);
            s += std::hypot(A[o], A[o + 1]);
        }
        best = std::max(best, s);
    }
    return best;
}

static void *prop_ptr(qjx_ctx *c, int b, int lvl)
{
    return (char *)c->props + ((size_t)(b * c->D + lvl)) * c->prop_elems * c->pesize;
}

static void band_of(qjx_ctx *c, int b, int lvl);

static void *prop_lo_ptr(qjx_ctx *c, int b, int lvl)
{
    if (!c->props_lo) return nullptr;
    return (char *)c->props_lo + ((size_t)(b * c->D + lvl)) * c->prop_elems * c->pesize;
}

// Band table of propagator (b, lvl) on the engine stream (MFMA kernels only read it).
static void band_of(qjx_ctx *c, int b, int lvl)
{
    const int g = b * c->D + lvl;
    qjx::launch_band(c->prec, c->cp, prop_ptr(c, b, lvl), prop_lo_ptr(c, b, lvl), c->L, c->dp,
                     c->dense || c->gemm_kind != qjx::kGemmMfma, c->krange + (size_t)2 * g * (c->L / 128),
                     c->kfrag + (size_t)qjx::kFragWords * g * (c->L / 128), c->stream);
}

extern "C" const char *qjx_last_error(void) { return g_err.c_str(); }

namespace qjx {
void set_last_error(const std::string &msg) { g_err = msg; }
}

extern "C" int qjx_dim(const qjx_ctx *c) { return c ? c->d : -1; }

static int upload(void **dst, const void *src, size_t bytes)
{
    HIP_TRY(hipMalloc(dst, bytes));
    HIP_TRY(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
    return QJX_OK;
}

static int env_gemm_kind()
{
    if (const char *env = std::getenv("QJX_GEMM"))
        if (std::string(env) == "fma") return qjx::kGemmFma;
    return qjx::kGemmMfma;
}

extern "C" int qjx_ctx_create(const qjx_model_desc *m, qjx_ctx **out)
{
    return qjx::ctx_create_kind(m, env_gemm_kind(), out);
}

int qjx::ctx_create_kind(const qjx_model_desc *m, int gemm_kind, qjx_ctx **out)
{
    if (!m || !out) return fail(QJX_ERR_ARG, "null argument");
    *out = nullptr;
    if (m->N < 1 || m->N + 1 > 2048) return fail(QJX_ERR_ARG, "N must be in [1, 2047]");
    if (m->ex_deep < 1 || m->ex_deep > qjx::kMaxGroups / 2 || m->ex_deep > 30)
        return fail(QJX_ERR_ARG, "ex_deep must be in [1, 30]");
    const int S = m->steps_per_period ? m->steps_per_period : 4;
    if (S < 2 || (S % 2)) return fail(QJX_ERR_ARG, "steps_per_period must be even and >= 2");
    if (m->precision != QJX_FP64 && m->precision != QJX_FP32 && m->precision != QJX_BF16X3)
        return fail(QJX_ERR_ARG, "unknown precision form");
    if (!(m->drv_freq > 0.0)) return fail(QJX_ERR_ARG, "dimer_drv_freq must be > 0");
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count < 1)
        return fail(QJX_ERR_HIP, "no HIP device available");

    qjx_ctx *c = new qjx_ctx;
    c->model = *m;
    c->model.steps_per_period = S;
    c->d = m->N + 1;
    c->dp = (c->d + 63) / 64 * 64;
    c->L = 2 * c->dp;
    c->D = m->ex_deep;
    c->S = S;
    c->G = 2 * c->D;
    c->gemm_kind = gemm_kind;
    c->R = c->L / qjx::gemm_bm(c->gemm_kind);
    c->prec = m->precision;
    c->esize = c->prec == QJX_FP64 ? 8 : 4;
    c->pesize = c->prec == QJX_FP64 ? 8 : c->prec == QJX_FP32 ? 4 : 2;
    c->cp = qjx::cp_layout(c->prec, c->gemm_kind);
    c->prop_elems = (size_t)(c->cp ? c->dp : c->L) * c->L;
    if (c->prec == QJX_BF16X3 && c->gemm_kind != qjx::kGemmMfma) {
        qjx_ctx_destroy(c);
        return fail(QJX_ERR_ARG, "the 3xBF16 form needs the matrix-core GEMM");
    }
    c->T = 2.0 * kPi / m->drv_freq;
    c->dt.resize(c->D);
    double dt = c->T / double(S);
    for (int i = 0; i < c->D; i++) {
        c->dt[i] = dt;
        dt *= 0.5;
    }
    build_model(c);
    int rc = QJX_OK;
    auto bail = [&](int code) {
        qjx_ctx_destroy(c);
        return code;
    };
    if (hipStreamCreate(&c->stream) != hipSuccess)
        return bail(fail(QJX_ERR_HIP, "hipStreamCreate failed"));
    if ((rc = upload((void **)&c->d_dt, c->dt.data(), sizeof(double) * c->D))) return bail(rc);
    if ((rc = upload((void **)&c->d_Ldiag, c->Ldiag.data(), sizeof(double) * c->d))) return bail(rc);
    if ((rc = upload((void **)&c->d_Loff, c->Loff.data(), sizeof(double) * c->Loff.size()))) return bail(rc);
    if ((rc = upload((void **)&c->d_HdiagE, c->HdiagE.data(), sizeof(double) * c->d))) return bail(rc);
    if ((rc = upload((void **)&c->d_Hoff, c->Hoff.data(), sizeof(double) * c->Hoff.size()))) return bail(rc);
    const size_t propbytes = (size_t)2 * c->D * c->prop_elems * c->pesize;
    if (hipMalloc(&c->props, propbytes) != hipSuccess)
        return bail(fail(QJX_ERR_NOMEM, "propagator tree does not fit in device memory"));
    if (c->prec == QJX_BF16X3 && hipMalloc(&c->props_lo, propbytes) != hipSuccess)
        return bail(fail(QJX_ERR_NOMEM, "propagator tree does not fit in device memory"));
    if (hipMalloc((void **)&c->krange, (size_t)2 * c->G * (c->L / 128) * sizeof(int)) != hipSuccess ||
        hipMalloc((void **)&c->kfrag, (size_t)qjx::kFragWords * c->G * (c->L / 128) * sizeof(unsigned)) != hipSuccess)
        return bail(fail(QJX_ERR_NOMEM, "band table does not fit in device memory"));
    if (const char *env = std::getenv("QJX_DENSE")) c->dense = std::string(env) == "1";

    // Propagator tree (qj_proc.cpp:4-120): U[b][i] = exp(Hqj[b] * dt_i).
    const size_t nn2 = (size_t)2 * c->d * c->d;
    double *dA = nullptr, *work = nullptr, *U = nullptr;
    if (hipMalloc((void **)&dA, nn2 * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&work, 6 * nn2 * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&U, 2 * nn2 * sizeof(double)) != hipSuccess) {
        dev_free(dA);
        dev_free(work);
        dev_free(U);
        return bail(fail(QJX_ERR_NOMEM, "propagator build workspace does not fit"));
    }
    bool direct = false;
    if (const char *env = std::getenv("QJX_EXPM_DIRECT")) direct = std::string(env) == "1";
    for (int b = 0; b < 2 && rc == QJX_OK; b++) {
        const double a1 = norm1_host(c->Hqj[b], c->d);
        // ordered on the engine stream: the previous branch's expm still reads dA
        if (hipMemcpyAsync(dA, c->Hqj[b].data(), nn2 * sizeof(double), hipMemcpyHostToDevice, c->stream) !=
                hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess) {
            rc = fail(QJX_ERR_HIP, "upload of H_eff failed");
            break;
        }
        // Deepest level first.  Where level i needs s_i >= 1 squarings, level i + 1
        // (dt_i / 2) needs s_i - 1 and starts from the same scaled argument
        // A dt_i / 2^s_i, so its result is exactly the scaling step's value after
        // s_i - 1 squarings: one more squaring gives level i bit for bit as the
        // full scaling and squaring would, at the cost of one product instead of
        // 7 + s_i.  Levels with s_i = 0 take the Taylor polynomial directly.
        // QJX_EXPM_DIRECT=1 runs the full scaling and squaring at every level.
        double *Ucur = U, *Uprev = U + nn2;
        for (int i = c->D - 1; i >= 0 && rc == QJX_OK; i--) {
            const int si = expm_squarings(a1, c->dt[i]);
            if (!direct && i + 1 < c->D && si >= 1 && expm_squarings(a1, c->dt[i + 1]) == si - 1) {
                qjx::launch_zgemm(c->d, Uprev, Uprev, Ucur, c->stream);
                if (hipGetLastError() != hipSuccess) rc = fail(QJX_ERR_HIP, "propagator squaring failed");
            } else {
                rc = device_expm(c, dA, a1, c->dt[i], Ucur, work);
            }
            if (rc == QJX_OK) {
                qjx::launch_pack_propagator(c->prec, c->cp, c->d, c->dp, Ucur, prop_ptr(c, b, i),
                                            prop_lo_ptr(c, b, i), c->stream);
                band_of(c, b, i);
            }
            std::swap(Ucur, Uprev);
        }
    }
    if (rc == QJX_OK && hipStreamSynchronize(c->stream) != hipSuccess)
        rc = fail(QJX_ERR_HIP, "propagator build failed");
    dev_free(dA);
    dev_free(work);
    dev_free(U);
    if (rc != QJX_OK) return bail(rc);
    *out = c;
    return QJX_OK;
}

// ---------------------------------------------------------------------------
// Runge-Kutta propagation (prop_id 1): RKPropagateBehavior::init_prop_data ->
// init_rk (propagator.cpp:66-69, rk_proc.cpp:56-63); the model matrices of
// init_hamiltonians_qj (newdel.cpp:2285-2365) as band coefficients.
// ---------------------------------------------------------------------------

// H_part = H - i/2 gamma' L^dag L (newdel.cpp:2327-2352) on its band |i - j| <= 2,
// H_drv and L, in the RK kernel's lane-major layout: entry i = l E + j of array a
// at (a E + j) P + l (P lanes per trajectory), zero beyond d.  (L^dag L)_ij sums L_ki L_kj in ascending k
// as build_model does, so H_part(i, j) and H_part(j, i) are the same number.
static void build_rk_bands(const qjx_ctx *c, int E, int P, std::vector<double> &out)
{
    const int d = c->d;
    const double g = c->model.diss_gamma / double(d - 1);
    auto Lel = [&](int i, int j) -> double {
        if (i == j) return c->Ldiag[i];
        if (i == j + 1) return -c->Loff[j];
        if (j == i + 1) return c->Loff[i];
        return 0.0;
    };
    auto hpart = [&](int i, int j, double *re, double *im) {
        double s = 0.0;
        for (int k = std::max(0, std::max(i, j) - 1); k <= std::min(d - 1, std::min(i, j) + 1); k++)
            s += Lel(k, i) * Lel(k, j);
        const double ldl = g * s;
        double h = 0.0;
        if (i == j) h = c->Hdiag[i];
        else if (i == j + 1) h = c->Hoff[j];
        else if (j == i + 1) h = c->Hoff[i];
        *re = h + 0.5 * 0.0;        // Im(diss_part) = 0
        *im = 0.0 - 0.5 * ldl;
    };
    out.assign((size_t)qjx::kRkBands * E * P, 0.0);
    auto at = [&](int a, int i) -> double & { return out[((size_t)a * E + i % E) * P + i / E]; };
    for (int i = 0; i < d; i++) {
        double re, im;
        hpart(i, i, &re, &im);
        at(qjx::kRkDre, i) = re;
        at(qjx::kRkDim, i) = im;
        at(qjx::kRkDrv, i) = c->Hdrv[i];
        at(qjx::kRkLd, i) = Lel(i, i);
        if (i + 1 < d) {
            hpart(i, i + 1, &re, &im);
            at(qjx::kRkO1re, i) = re;
            at(qjx::kRkO1im, i) = im;
            at(qjx::kRkLp, i) = Lel(i, i + 1);
        }
        if (i + 2 < d) {
            hpart(i, i + 2, &re, &im);
            at(qjx::kRkO2im, i) = im;
        }
        if (i >= 1) at(qjx::kRkLm, i) = Lel(i, i - 1);
    }
}

extern "C" int qjx_ctx_create_rk(const qjx_model_desc *m, const qjx_rk_desc *rk, qjx_ctx **out)
{
    if (!m || !rk || !out) return fail(QJX_ERR_ARG, "null argument");
    *out = nullptr;
    if (m->N < 1 || m->N + 1 > qjx::kMaxDim) return fail(QJX_ERR_ARG, "N must be in [1, 2047]");
    if (m->precision != QJX_FP64)
        return fail(QJX_ERR_ARG, "the Runge-Kutta path is FP64 only (complex128, as the reference)");
    if (rk->rk_ns < 1) return fail(QJX_ERR_ARG, "rk_ns must be >= 1");
    if (!(m->drv_freq > 0.0)) return fail(QJX_ERR_ARG, "dimer_drv_freq must be > 0");
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count < 1)
        return fail(QJX_ERR_HIP, "no HIP device available");
    qjx_ctx *c = new qjx_ctx;
    c->model = *m;
    c->prop = 1;
    c->d = m->N + 1;
    c->dp = qjx::rk_group(c->d) * qjx::rk_lanes(c->d);   // whole lane chunks: the RK kernel reads and writes P E entries
    c->L = 2 * c->dp;
    c->D = 1;
    c->S = 2;
    c->G = 2;
    c->gemm_kind = qjx::kGemmMfma;
    c->R = c->L / qjx::gemm_bm(c->gemm_kind);
    c->prec = QJX_FP64;
    c->esize = c->pesize = 8;
    c->T = 2.0 * kPi / m->drv_freq;                     // newdel.cpp:31
    c->dt.assign(1, c->T);
    c->rk_ns = rk->rk_ns;
    c->rk_drv_type = rk->drv_type;
    c->rk_phase = rk->drv_phase;
    c->rk_step = c->T / double(rk->rk_ns);              // rk_proc.cpp:62
    build_model(c);
    auto bail = [&](int code) {
        qjx_ctx_destroy(c);
        return code;
    };
    int rc = QJX_OK;
    if (hipStreamCreate(&c->stream) != hipSuccess) return bail(fail(QJX_ERR_HIP, "hipStreamCreate failed"));
    if ((rc = upload((void **)&c->d_dt, c->dt.data(), sizeof(double)))) return bail(rc);
    if ((rc = upload((void **)&c->d_Ldiag, c->Ldiag.data(), sizeof(double) * c->d))) return bail(rc);
    if ((rc = upload((void **)&c->d_Loff, c->Loff.data(), sizeof(double) * c->Loff.size()))) return bail(rc);
    if ((rc = upload((void **)&c->d_HdiagE, c->HdiagE.data(), sizeof(double) * c->d))) return bail(rc);
    if ((rc = upload((void **)&c->d_Hoff, c->Hoff.data(), sizeof(double) * c->Hoff.size()))) return bail(rc);
    std::vector<double> bands;
    build_rk_bands(c, qjx::rk_lanes(c->d), qjx::rk_group(c->d), bands);
    if ((rc = upload((void **)&c->d_rkbands, bands.data(), sizeof(double) * bands.size()))) return bail(rc);
    *out = c;
    return QJX_OK;
}

// Advance every trajectory by n RK steps (rk_period, core.cpp:439-450, over all
// trajectories).  With logging on, one period per launch and a drain after each;
// otherwise launches of bounded work.
static int check_err(qjx_ctx *c);
static int drain_log(qjx_ctx *c);

static int rk_run(qjx_ctx *c, long long n, int is_obs)
{
    const bool logging = is_obs && c->ens.jump > 0;
    const double work = std::max(1.0, double(c->M) * double(c->d));
    const long long chunk = std::min(logging ? (long long)c->rk_ns : std::max(1LL, (long long)(2.0e10 / work)),
                                     1LL << 22);
    const long long tab_steps = std::min(chunk, n);
    if (tab_steps > c->rk_tab_steps) {
        dev_free(c->d_rktab);
        c->rk_tab_steps = 0;
        HIP_TRY(hipMalloc((void **)&c->d_rktab, (size_t)tab_steps * 3 * sizeof(double)));
        c->rk_tab_steps = tab_steps;
    }
    hipEvent_t ev[2] = {nullptr, nullptr};
    if (c->profiling) {
        HIP_TRY(hipEventCreate(&ev[0]));
        HIP_TRY(hipEventCreate(&ev[1]));
    }
    int rc = QJX_OK;
    while (n > 0 && rc == QJX_OK) {
        const long long k = std::min(n, chunk);
        qjx::RkArgs a{};
        a.bands = c->d_rkbands;
        a.tab = c->d_rktab;
        a.slot0 = c->slot[0];
        a.slot1 = c->slot[1];
        a.st = c->st;
        a.log = c->log;
        a.err = c->err;
        a.counters = c->counters;
        a.stride = c->stride;
        a.pos0 = c->rk_pos;
        a.T = c->T;
        a.rk_step = c->rk_step;
        a.prm_E = c->model.prm_E;
        a.ampl = c->model.drv_ampl;
        a.freq = c->model.drv_freq;
        a.phase = c->rk_phase;
        a.M = c->M;
        a.d = c->d;
        a.dp = c->dp;
        a.L = c->L;
        a.nsteps = (int)k;
        a.rk_ns = c->rk_ns;
        a.drv_type = c->rk_drv_type;
        a.jump = c->ens.jump;
        a.is_obs = is_obs;
        if (c->profiling && hipEventRecord(ev[0], c->stream) != hipSuccess) rc = fail(QJX_ERR_HIP, "event record");
        const hipError_t e = qjx::launch_rk(a, c->stream);
        if (e != hipSuccess) {
            rc = fail(QJX_ERR_HIP, std::string("RK launch: ") + hipGetErrorString(e));
            break;
        }
        if (c->profiling && hipEventRecord(ev[1], c->stream) != hipSuccess) rc = fail(QJX_ERR_HIP, "event record");
        c->rk_launches++;
        c->rk_pos += k;
        n -= k;
        if (rc == QJX_OK) rc = check_err(c);      // synchronises
        if (rc == QJX_OK && c->profiling) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, ev[0], ev[1]) == hipSuccess) c->rk_ms += ms;
        }
        if (rc == QJX_OK && logging) rc = drain_log(c);
    }
    for (auto &x : ev)
        if (x) (void)hipEventDestroy(x);
    return rc;
}

static void free_ensemble(qjx_ctx *c)
{
    for (int i = 0; i < 2; i++) {
        if (c->slot[i]) (void)hipFree(c->slot[i]);
        c->slot[i] = nullptr;
        dev_free(c->lists[i]);
        dev_free(c->cnts[i]);
    }
    dev_free(c->st);
    dev_free(c->normpart);
    dev_free(c->err);
    dev_free(c->counters);
    dev_free(c->log.time);
    dev_free(c->log.norm);
    dev_free(c->log.eta);
    dev_free(c->log.type);
    dev_free(c->mean_start);
    dev_free(c->obs);
    dev_free(c->adr);
    dev_free(c->adr_avg);
    dev_free(c->xfer);
    if (c->runner) qjx::runner_destroy(c->runner);
    c->runner = nullptr;
    c->have_ens = false;
}

extern "C" void qjx_ctx_destroy(qjx_ctx *c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_ensemble(c);
    if (c->props) (void)hipFree(c->props);
    if (c->props_lo) (void)hipFree(c->props_lo);
    dev_free(c->krange);
    dev_free(c->kfrag);
    dev_free(c->d_dt);
    dev_free(c->d_Ldiag);
    dev_free(c->d_Loff);
    dev_free(c->d_HdiagE);
    dev_free(c->d_Hoff);
    dev_free(c->d_rkbands);
    dev_free(c->d_rktab);
    if (c->robs_R) (void)hipFree(c->robs_R);
    if (c->robs_scratch) (void)hipFree(c->robs_scratch);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

static int alloc_log(qjx_ctx *c, int cap)
{
    dev_free(c->log.time);
    dev_free(c->log.norm);
    dev_free(c->log.eta);
    dev_free(c->log.type);
    c->log.cap = 0;
    if (c->ens.jump <= 0 || cap <= 0) return QJX_OK;
    const size_t n = (size_t)c->M * cap;
    HIP_TRY(hipMalloc((void **)&c->log.time, n * sizeof(double)));
    HIP_TRY(hipMalloc((void **)&c->log.norm, n * sizeof(double)));
    HIP_TRY(hipMalloc((void **)&c->log.eta, n * sizeof(double)));
    HIP_TRY(hipMalloc((void **)&c->log.type, n * sizeof(int)));
    c->log.cap = cap;
    return QJX_OK;
}

static void build_desc(qjx_ctx *c)
{
    qjx::Ens &E = c->desc;
    E.props = c->props;
    E.props_lo = c->props_lo;
    E.krange = c->krange;
    E.kfrag = c->kfrag;
    for (int i = 0; i < 2; i++) {
        E.slot[i] = c->slot[i];
        E.list[i] = c->lists[i];
        E.cnt[i] = c->cnts[i];
    }
    E.st = c->st;
    E.normpart = c->normpart;
    E.dt = c->d_dt;
    E.Ldiag = c->d_Ldiag;
    E.Loff = c->d_Loff;
    E.log = c->log;
    E.err = c->err;
    E.counters = c->counters;
    E.stride = c->stride;
    // reduced-precision forms: ||psi||^2 may exceed 1 by the GEMM's rounding
    E.norm_tol = c->prec == QJX_FP64 ? 0.0 : c->prec == QJX_FP32 ? 1e-4 : 1e-3;
    E.d = c->d;
    E.dp = c->dp;
    E.L = c->L;
    E.D = c->D;
    E.S = c->S;
    E.M = c->M;
    E.G = c->G;
    E.R = c->R;
    E.is_obs = 0;
    E.jump = c->ens.jump;
}

static int ensure_runner(qjx_ctx *c)
{
    if (c->runner) return QJX_OK;
    int rc = qjx::runner_create(c->prec, c->gemm_kind, c->stream, c->M, &c->runner);
    if (rc) return rc;
    c->runner->profiling = c->profiling;
    if (const char *env = std::getenv("QJX_GRAPH"))
        if (std::string(env) == "0") c->runner->graphs = false;
    return QJX_OK;
}

extern "C" int qjx_ensemble_init(qjx_ctx *c, const qjx_ensemble_desc *e)
{
    if (!c || !e) return fail(QJX_ERR_ARG, "null argument");
    if (e->num_trajectories < 1 || e->num_trajectories > (1 << 30))
        return fail(QJX_ERR_ARG, "num_trajectories must be in [1, 2^30]");
    if (e->mns < 1) return fail(QJX_ERR_ARG, "mns must be >= 1");
    if (e->start_type == 0 && (e->start_state < 0 || e->start_state >= c->d))
        return fail(QJX_ERR_ARG, "start_state out of range");
    free_ensemble(c);
    c->ens = *e;
    c->M = e->num_trajectories;
    c->stride = mcg_pow_host(kMcgA, (uint64_t)e->mns);
    const size_t M = c->M;
    const size_t slotbytes = M * c->L * c->esize;
    HIP_TRY(hipMalloc(&c->slot[0], slotbytes));
    HIP_TRY(hipMalloc(&c->slot[1], slotbytes));
    HIP_TRY(hipMalloc((void **)&c->st, M * sizeof(qjx::TrajState)));
    for (int i = 0; i < 2; i++) {
        HIP_TRY(hipMalloc((void **)&c->lists[i], (size_t)c->G * M * sizeof(int)));
        HIP_TRY(hipMalloc((void **)&c->cnts[i], qjx::kMaxGroups * sizeof(int)));
        HIP_TRY(hipMemset(c->cnts[i], 0, qjx::kMaxGroups * sizeof(int)));
    }
    HIP_TRY(hipMalloc((void **)&c->normpart, (size_t)c->R * M * sizeof(double)));
    HIP_TRY(hipMalloc((void **)&c->err, 4 * sizeof(int)));
    HIP_TRY(hipMemset(c->err, 0, 4 * sizeof(int)));
    HIP_TRY(hipMalloc((void **)&c->counters, 4 * sizeof(unsigned long long)));
    HIP_TRY(hipMemset(c->counters, 0, 4 * sizeof(unsigned long long)));
    HIP_TRY(hipMalloc((void **)&c->mean_start, M * sizeof(double)));
    HIP_TRY(hipMemset(c->mean_start, 0, M * sizeof(double)));
    HIP_TRY(hipMalloc((void **)&c->obs, 5 * M * sizeof(double)));
    const int cap = e->log_capacity > 0 ? e->log_capacity : 256;
    int rc = alloc_log(c, cap);
    if (rc) return rc;
    c->jt.assign(M, {});
    c->jn.assign(M, {});
    c->je.assign(M, {});
    c->jty.assign(M, {});
    qjx::launch_init_ensemble(c->prec, c->slot[0], c->slot[1], c->st, c->M, c->d, c->dp, c->L,
                              e->seed, e->start_type, e->start_state, c->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    build_desc(c);
    c->have_ens = true;
    return QJX_OK;
}

static int check_err(qjx_ctx *c)
{
    int h[4];
    HIP_TRY(hipMemcpy(h, c->err, sizeof h, hipMemcpyDeviceToHost));
    if (h[0] == 2) {
        char buf[200];
        std::snprintf(buf, sizeof buf, "trajectory %d: non-finite norm or endless crossings in one RK step", h[1]);
        return fail(QJX_ERR_NORM, buf);
    }
    if (h[0]) {
        const unsigned long long bits = (unsigned long long)(unsigned)h[2] | ((unsigned long long)(unsigned)h[3] << 32);
        double n;
        std::memcpy(&n, &bits, sizeof n);
        char buf[200];
        std::snprintf(buf, sizeof buf, "norm more than 1: trajectory %d, ||psi||^2 = %.17g", h[1], n);
        return fail(QJX_ERR_NORM, buf);
    }
    return QJX_OK;
}

// Copy the device jump log into the host vectors and reset the device counters.
static int drain_log(qjx_ctx *c)
{
    if (c->log.cap <= 0) return QJX_OK;
    const size_t M = c->M, cap = c->log.cap;
    std::vector<qjx::TrajState> hs(M);
    HIP_TRY(hipMemcpy(hs.data(), c->st, M * sizeof(qjx::TrajState), hipMemcpyDeviceToHost));
    bool any = false, over = false;
    for (size_t tr = 0; tr < M; tr++) {
        any = any || hs[tr].logn > 0;
        over = over || hs[tr].logn > (int)cap;
    }
    if (!any) return QJX_OK;
    std::vector<double> t(M * cap), n(M * cap), e(M * cap);
    std::vector<int> ty(M * cap);
    HIP_TRY(hipMemcpy(t.data(), c->log.time, t.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(n.data(), c->log.norm, n.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(e.data(), c->log.eta, e.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(ty.data(), c->log.type, ty.size() * sizeof(int), hipMemcpyDeviceToHost));
    for (size_t tr = 0; tr < M; tr++) {
        const int k = std::min(hs[tr].logn, (int)cap);
        for (int i = 0; i < k; i++) {
            c->jt[tr].push_back(t[tr * cap + i]);
            c->jn[tr].push_back(n[tr * cap + i]);
            c->je[tr].push_back(e[tr * cap + i]);
            c->jty[tr].push_back(ty[tr * cap + i]);
        }
    }
    qjx::launch_reset_logn(c->st, c->M, c->stream);
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (over) return fail(QJX_ERR_LOG, "jump log overflow: raise log_capacity or drain more often");
    return QJX_OK;
}

// Advance every trajectory by n top-level steps: seed the lists, then run
// iteration chunks until no trajectory has steps left.
static long long steps_per_period(const qjx_ctx *c) { return c->prop == 1 ? c->rk_ns : c->S; }

static int run_steps(qjx_ctx *c, long long n, int is_obs)
{
    if (n <= 0) return QJX_OK;
    if (c->prop == 1) return rk_run(c, n, is_obs);
    if (n > (1LL << 30)) return fail(QJX_ERR_ARG, "too many steps in one call");
    int rc = ensure_runner(c);
    if (rc) return rc;
    qjx::launch_seed(c->st, c->M, (int)n, c->lists[0], c->cnts[0], c->D, c->stream);
    HIP_TRY(hipGetLastError());
    c->desc.is_obs = is_obs;
    rc = qjx::runner_set_table(c->runner, &c->desc, 1);
    if (rc) return rc;
    rc = qjx::runner_run(c->runner, nullptr);
    if (rc) return rc;
    return check_err(c);
}

extern "C" int qjx_trans(qjx_ctx *c, int num_tp_periods)
{
    if (!c || !c->have_ens) return fail(QJX_ERR_ARG, "ensemble not initialised");
    if (num_tp_periods < 0) return fail(QJX_ERR_ARG, "negative period count");
    qjx::launch_draw_eta(c->st, c->M, c->stride, c->stream);
    // RK: transient periods are numbered from 0, and so are the observed ones
    // (obser_process starts at dump_periods[0] = 0, experiment.cpp:476-488)
    c->rk_pos = 0;
    int rc = run_steps(c, (long long)num_tp_periods * steps_per_period(c), 0);
    if (rc) return rc;
    c->rk_pos = 0;
    qjx::launch_reset_times(c->st, c->M, c->stream);
    HIP_TRY(hipStreamSynchronize(c->stream));
    return QJX_OK;
}

extern "C" int qjx_run_periods(qjx_ctx *c, int n, int is_obs)
{
    if (!c || !c->have_ens) return fail(QJX_ERR_ARG, "ensemble not initialised");
    if (n < 0) return fail(QJX_ERR_ARG, "negative period count");
    int rc = run_steps(c, (long long)n * steps_per_period(c), is_obs);
    if (rc) return rc;
    if (is_obs && c->ens.jump > 0) return drain_log(c);
    return QJX_OK;
}

extern "C" int qjx_run_steps(qjx_ctx *c, int n, int is_obs)
{
    if (!c || !c->have_ens) return fail(QJX_ERR_ARG, "ensemble not initialised");
    if (n < 0) return fail(QJX_ERR_ARG, "negative step count");
    int rc = run_steps(c, n, is_obs);
    if (rc) return rc;
    if (is_obs && c->ens.jump > 0) return drain_log(c);
    return QJX_OK;
}

extern "C" int qjx_chars(qjx_ctx *c, int is_start, double *norm, double *mean, double *dispersion,
                         double *m2, double *energy, double *adr_avg, double *adr)
{
    if (!c || !c->have_ens) return fail(QJX_ERR_ARG, "ensemble not initialised");
    const size_t M = c->M;
    const bool need_adr = adr_avg || adr;
    if (need_adr && !c->adr) {
        HIP_TRY(hipMalloc((void **)&c->adr, M * c->d * sizeof(double)));
        HIP_TRY(hipMalloc((void **)&c->adr_avg, (size_t)(1 + 64) * c->d * sizeof(double)));
    }
    double *o = c->obs;
    qjx::launch_chars(c->prec, c->slot[0], c->slot[1], c->st, c->M, c->d, c->dp, c->L, c->d_HdiagE,
                      c->d_Hoff, is_start, c->mean_start, o, o + M, o + 2 * M, o + 3 * M, o + 4 * M,
                      need_adr ? c->adr : nullptr, c->stream);
    if (adr_avg) qjx::launch_adr_avg(c->adr, c->M, c->d, c->adr_avg, c->stream);
    HIP_TRY(hipGetLastError());
    double *dst[5] = {norm, mean, dispersion, m2, energy};
    for (int i = 0; i < 5; i++)
        if (dst[i]) HIP_TRY(hipMemcpyAsync(dst[i], o + i * M, M * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (adr_avg) HIP_TRY(hipMemcpyAsync(adr_avg, c->adr_avg, c->d * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (adr) HIP_TRY(hipMemcpyAsync(adr, c->adr, M * c->d * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return QJX_OK;
}

extern "C" int qjx_density(qjx_ctx *c, double *rho, double *purity)
{
    if (!c || !c->have_ens) return fail(QJX_ERR_ARG, "ensemble not initialised");
    const int splits = qjx::density_splits(c->M, c->L);
    const size_t nrho = 2 * (size_t)c->d * c->d;
    double *w = nullptr, *part = nullptr, *drho = nullptr;
    auto release = [&]() {
        dev_free(w);
        dev_free(part);
        dev_free(drho);
    };
    if (hipMalloc((void **)&w, ((size_t)c->M + 1) * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&part, (size_t)splits * c->L * c->L * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&drho, nrho * sizeof(double)) != hipSuccess) {
        release();
        return fail(QJX_ERR_NOMEM, "density scratch does not fit");
    }
    qjx::launch_density(c->prec, c->slot[0], c->slot[1], c->st, c->M, c->d, c->dp, c->L, w, part, splits,
                        drho, w + c->M, c->stream);
    int rc = QJX_OK;
    if (hipGetLastError() != hipSuccess ||
        (rho && hipMemcpyAsync(rho, drho, nrho * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess) ||
        (purity && hipMemcpyAsync(purity, w + c->M, sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess) ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        rc = fail(QJX_ERR_HIP, "density launch failed");
    release();
    return rc;
}

extern "C" int qjx_get_state(qjx_ctx *c, double *psi, double *eta, double *t, int *njumps, uint64_t *rng)
{
    if (!c || !c->have_ens) return fail(QJX_ERR_ARG, "ensemble not initialised");
    const size_t M = c->M;
    if (psi) {
        if (!c->xfer) HIP_TRY(hipMalloc((void **)&c->xfer, M * c->d * 2 * sizeof(double)));
        qjx::launch_export_state(c->prec, c->slot[0], c->slot[1], c->st, c->M, c->d, c->dp, c->L, c->xfer, c->stream);
        HIP_TRY(hipMemcpyAsync(psi, c->xfer, M * c->d * 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    }
    std::vector<qjx::TrajState> hs(M);
    HIP_TRY(hipMemcpyAsync(hs.data(), c->st, M * sizeof(qjx::TrajState), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (size_t tr = 0; tr < M; tr++) {
        if (eta) eta[tr] = hs[tr].eta;
        if (t) t[tr] = hs[tr].t;
        if (njumps) njumps[tr] = hs[tr].njumps;
        if (rng) rng[tr] = hs[tr].rng;
    }
    return QJX_OK;
}

extern "C" int qjx_set_state(qjx_ctx *c, const double *psi, const double *eta, const double *t,
                             const uint64_t *rng)
{
    if (!c || !c->have_ens) return fail(QJX_ERR_ARG, "ensemble not initialised");
    const size_t M = c->M;
    std::vector<qjx::TrajState> hs(M);
    HIP_TRY(hipMemcpy(hs.data(), c->st, M * sizeof(qjx::TrajState), hipMemcpyDeviceToHost));
    for (size_t tr = 0; tr < M; tr++) {
        if (eta) hs[tr].eta = eta[tr];
        if (t) hs[tr].t = t[tr];
        if (rng) hs[tr].rng = rng[tr];
    }
    HIP_TRY(hipMemcpy(c->st, hs.data(), M * sizeof(qjx::TrajState), hipMemcpyHostToDevice));
    if (psi) {
        if (!c->xfer) HIP_TRY(hipMalloc((void **)&c->xfer, M * c->d * 2 * sizeof(double)));
        HIP_TRY(hipMemcpy(c->xfer, psi, M * c->d * 2 * sizeof(double), hipMemcpyHostToDevice));
        qjx::launch_import_state(c->prec, c->slot[0], c->slot[1], c->st, c->M, c->d, c->dp, c->L, c->xfer, c->stream);
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return QJX_OK;
}

extern "C" int qjx_get_jumps(qjx_ctx *c, int tr, double *times, int *types, double *norms, double *etas, int cap)
{
    if (!c || !c->have_ens) return fail(QJX_ERR_ARG, "ensemble not initialised");
    if (tr < 0 || tr >= c->M) return fail(QJX_ERR_ARG, "trajectory out of range");
    const int n = (int)c->jt[tr].size();
    for (int i = 0; i < n && i < cap; i++) {
        if (times) times[i] = c->jt[tr][i];
        if (types) types[i] = c->jty[tr][i];
        if (norms) norms[i] = c->jn[tr][i];
        if (etas) etas[i] = c->je[tr][i];
    }
    return n;
}

// Waiting-time distribution of the logged jumps, binned as the reference's
// post-processing does (plot/mpipks/qjx/dimer/times_between_jumps.m:41-131): log-spaced
// bins [b0 10^((i-1)/n), b0 10^(i/n)), bin of x = floor((log10 x - log10 b0) nb /
// (log10 b_end - log10 b0 + eps)) + 1, x outside [b0, b_end] counted apart, the pdf
// normalised by the in-range count and the bin widths.
extern "C" int qjx_jump_interval_pdf(qjx_ctx *c, double bin_begin, int num_decades, int bins_per_decade,
                                     double *pdf, int64_t *counts, int64_t *outside)
{
    if (!c || !c->have_ens) return fail(QJX_ERR_ARG, "ensemble not initialised");
    if (!(bin_begin > 0.0) || num_decades < 1 || bins_per_decade < 1)
        return fail(QJX_ERR_ARG, "bin_begin must be > 0, num_decades and bins_per_decade >= 1");
    const int nb = num_decades * bins_per_decade;
    const double bin_end = bin_begin * std::pow(10.0, num_decades);
    const double eps = 2.220446049250313e-16;   // MATLAB eps
    const double lb = std::log10(bin_begin), span = std::log10(bin_end) - lb + eps;
    std::vector<int64_t> cnt(nb, 0);
    int64_t out = 0, total = 0;
    for (const auto &jt : c->jt)
        for (size_t i = 1; i < jt.size(); i++) {
            const double x = jt[i] - jt[i - 1];
            if (x >= bin_begin && x <= bin_end) {
                const int b = (int)std::floor((std::log10(x) - lb) * nb / span);
                cnt[std::min(std::max(b, 0), nb - 1)]++;
                total++;
            } else {
                out++;
            }
        }
    for (int i = 0; i < nb; i++) {
        const double lo = bin_begin * std::pow(10.0, double(i) / bins_per_decade);
        const double hi = bin_begin * std::pow(10.0, double(i + 1) / bins_per_decade);
        if (pdf) pdf[i] = total ? double(cnt[i]) / (double(total) * (hi - lo)) : 0.0;
        if (counts) counts[i] = cnt[i];
    }
    if (outside) *outside = out;
    return QJX_OK;
}

extern "C" int qjx_get_propagator(qjx_ctx *c, int b, int lvl, double *U)
{
    if (c && c->prop == 1) return fail(QJX_ERR_ARG, "a Runge-Kutta context has no propagators");
    if (!c || !U || b < 0 || b > 1 || lvl < 0 || lvl >= c->D) return fail(QJX_ERR_ARG, "bad propagator index");
    std::vector<char> raw(c->prop_elems * c->pesize), rawlo;
    HIP_TRY(hipMemcpy(raw.data(), prop_ptr(c, b, lvl), raw.size(), hipMemcpyDeviceToHost));
    if (c->props_lo) {
        rawlo.resize(raw.size());
        HIP_TRY(hipMemcpy(rawlo.data(), prop_lo_ptr(c, b, lvl), rawlo.size(), hipMemcpyDeviceToHost));
    }
    auto bf = [](const char *p, size_t o) -> double {
        uint16_t h;
        std::memcpy(&h, p + 2 * o, 2);
        const uint32_t u = (uint32_t)h << 16;
        float f;
        std::memcpy(&f, &u, 4);
        return f;
    };
    auto el = [&](size_t o) -> double {
        if (c->pesize == 8) return ((const double *)raw.data())[o];
        if (c->pesize == 4) return ((const float *)raw.data())[o];
        return bf(raw.data(), o) + bf(rawlo.data(), o);
    };
    const int d = c->d, dp = c->dp, L = c->L;
    for (int i = 0; i < d; i++)
        for (int k = 0; k < d; k++) {
            // Re U_ik at (row i, col k); Im U_ik at (row i, col dp + k) in the
            // complex-pair layout, at (row dp + i, col k) in the block-real one
            U[2 * ((size_t)i * d + k)] = el((size_t)i * L + k);
            U[2 * ((size_t)i * d + k) + 1] = c->cp ? el((size_t)i * L + dp + k) : el((size_t)(dp + i) * L + k);
        }
    return QJX_OK;
}

extern "C" int qjx_set_propagator(qjx_ctx *c, int b, int lvl, const double *U)
{
    if (c && c->prop == 1) return fail(QJX_ERR_ARG, "a Runge-Kutta context has no propagators");
    if (!c || !U || b < 0 || b > 1 || lvl < 0 || lvl >= c->D) return fail(QJX_ERR_ARG, "bad propagator index");
    const size_t nn2 = (size_t)2 * c->d * c->d;
    double *dU = nullptr;
    HIP_TRY(hipMalloc((void **)&dU, nn2 * sizeof(double)));
    hipError_t e = hipMemcpy(dU, U, nn2 * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        qjx::launch_pack_propagator(c->prec, c->cp, c->d, c->dp, dU, prop_ptr(c, b, lvl), prop_lo_ptr(c, b, lvl),
                                    c->stream);
        band_of(c, b, lvl);
        e = hipStreamSynchronize(c->stream);
    }
    (void)hipFree(dU);
    if (e != hipSuccess) return fail(QJX_ERR_HIP, hipGetErrorString(e));
    return QJX_OK;
}

extern "C" int qjx_get_stats(qjx_ctx *c, qjx_stats *out)
{
    if (!c || !out) return fail(QJX_ERR_ARG, "null argument");
    unsigned long long h[3] = {0, 0, 0};
    if (c->counters) HIP_TRY(hipMemcpy(h, c->counters, sizeof h, hipMemcpyDeviceToHost));
    out->matvecs = (int64_t)h[0];
    out->iterations = c->runner ? c->runner->iterations : 0;
    out->jumps = (int64_t)h[1];
    out->gemm_flops = 8.0 * double(c->d) * double(c->d) * double(out->matvecs);
    out->gemm_band_flops = double(h[2]) * 128.0;
    out->gemm_ms = c->runner ? c->runner->gemm_ms : 0.0;
    out->decide_ms = c->runner ? c->runner->tail_ms + c->runner->plan_ms : 0.0;
    out->gemm_launches = c->runner ? c->runner->gemm_launches : 0;
    out->gemm_bn = c->runner ? c->runner->bn : 0;
    out->gemm_grid = c->runner ? c->runner->grid.gemm : 0;
    out->gemm_waves = c->runner ? c->runner->grid.waves : 0;
    if (c->prop == 1) {                 // RK: matvecs = right-hand-side evaluations
        out->iterations = c->rk_launches;
        out->gemm_flops = 0.0;
        out->gemm_band_flops = 0.0;
        out->gemm_ms = c->rk_ms;
        out->gemm_launches = c->rk_launches;
    }
    return QJX_OK;
}

extern "C" int qjx_mfma_probe_f32x16(const float *A, const float *B, float *D)
{
    if (!A || !B || !D) return fail(QJX_ERR_ARG, "null argument");
    float *buf = nullptr;   // A 64 | B 64 | D 256 floats
    HIP_TRY(hipMalloc((void **)&buf, (64 + 64 + 256) * sizeof(float)));
    hipError_t e = hipMemcpy(buf, A, 64 * sizeof(float), hipMemcpyHostToDevice);
    e = e == hipSuccess ? hipMemcpy(buf + 64, B, 64 * sizeof(float), hipMemcpyHostToDevice) : e;
    if (e == hipSuccess) {
        qjx::launch_mfma_probe16(buf, buf + 64, buf + 128, nullptr);
        e = hipGetLastError();
    }
    e = e == hipSuccess ? hipDeviceSynchronize() : e;
    e = e == hipSuccess ? hipMemcpy(D, buf + 128, 256 * sizeof(float), hipMemcpyDeviceToHost) : e;
    (void)hipFree(buf);
    if (e != hipSuccess) return fail(QJX_ERR_HIP, hipGetErrorString(e));
    return QJX_OK;
}

extern "C" int qjx_mfma_probe(const float *A32, const float *B32, float *D32, const double *A64,
                              const double *B64, double *D64, const float *Abf, const float *Bbf, float *Dbf)
{
    // one buffer: A32 64 | B32 64 | D32 1024 | Abf 512 | Bbf 512 | Dbf 1024 floats, then
    // A64 64 | B64 64 | D64 256 doubles
    const size_t nf = 64 + 64 + 1024 + 512 + 512 + 1024, nd = 64 + 64 + 256;
    char *buf = nullptr;
    HIP_TRY(hipMalloc((void **)&buf, nf * 4 + nd * 8));
    float *f = (float *)buf;
    double *dd = (double *)(buf + nf * 4);
    hipError_t e = hipSuccess;
    e = e == hipSuccess ? hipMemcpy(f, A32, 64 * 4, hipMemcpyHostToDevice) : e;
    e = e == hipSuccess ? hipMemcpy(f + 64, B32, 64 * 4, hipMemcpyHostToDevice) : e;
    e = e == hipSuccess ? hipMemcpy(f + 1152, Abf, 512 * 4, hipMemcpyHostToDevice) : e;
    e = e == hipSuccess ? hipMemcpy(f + 1664, Bbf, 512 * 4, hipMemcpyHostToDevice) : e;
    e = e == hipSuccess ? hipMemcpy(dd, A64, 64 * 8, hipMemcpyHostToDevice) : e;
    e = e == hipSuccess ? hipMemcpy(dd + 64, B64, 64 * 8, hipMemcpyHostToDevice) : e;
    if (e == hipSuccess) {
        qjx::launch_mfma_probe(f, f + 64, f + 128, dd, dd + 64, dd + 128, f + 1152, f + 1664, f + 2176, nullptr);
        e = hipGetLastError();
    }
    e = e == hipSuccess ? hipDeviceSynchronize() : e;
    e = e == hipSuccess ? hipMemcpy(D32, f + 128, 1024 * 4, hipMemcpyDeviceToHost) : e;
    e = e == hipSuccess ? hipMemcpy(Dbf, f + 2176, 1024 * 4, hipMemcpyDeviceToHost) : e;
    e = e == hipSuccess ? hipMemcpy(D64, dd + 128, 256 * 8, hipMemcpyDeviceToHost) : e;
    (void)hipFree(buf);
    if (e != hipSuccess) return fail(QJX_ERR_HIP, hipGetErrorString(e));
    return QJX_OK;
}

extern "C" int qjx_set_profiling(qjx_ctx *c, int on)
{
    if (!c) return fail(QJX_ERR_ARG, "null context");
    c->profiling = on != 0;
    c->rk_ms = 0.0;
    if (c->runner) {
        c->runner->profiling = c->profiling;
        c->runner->plan_ms = c->runner->gemm_ms = c->runner->tail_ms = 0.0;
        c->runner->gemm_launches = 0;
    }
    return QJX_OK;
}

// ---------------------------------------------------------------------------
// Batches of contexts sharing one runner
// ---------------------------------------------------------------------------
struct qjx_batch {
    std::vector<qjx_ctx *> ctx;
    hipStream_t stream = nullptr;
    qjx::Runner *runner = nullptr;
};

extern "C" void qjx_batch_destroy(qjx_batch *b)
{
    if (!b) return;
    qjx::runner_destroy(b->runner);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

extern "C" int qjx_batch_create(qjx_ctx *const *ctxs, int n, qjx_batch **out)
{
    if (!ctxs || !out) return fail(QJX_ERR_ARG, "null argument");
    *out = nullptr;
    if (n < 1 || n > qjx::kMaxEns) return fail(QJX_ERR_ARG, "a batch holds 1 to 64 contexts");
    for (int i = 0; i < n; i++)
        if (ctxs[i] && ctxs[i]->prop == 1)
            return fail(QJX_ERR_ARG, "Runge-Kutta contexts are not batched (no step GEMM to share)");
    long long total = 0;
    for (int i = 0; i < n; i++) {
        const qjx_ctx *c = ctxs[i];
        if (!c || !c->have_ens) return fail(QJX_ERR_ARG, "batch member without an initialised ensemble");
        if (c->prec != ctxs[0]->prec || c->gemm_kind != ctxs[0]->gemm_kind)
            return fail(QJX_ERR_ARG, "batch members must share the precision form");
        for (int j = 0; j < i; j++)
            if (ctxs[j] == c) return fail(QJX_ERR_ARG, "a context appears twice in the batch");
        total += c->M;
    }
    if (total > (1LL << 30)) return fail(QJX_ERR_ARG, "too many trajectories in one batch");
    qjx_batch *b = new qjx_batch;
    b->ctx.assign(ctxs, ctxs + n);
    if (hipStreamCreate(&b->stream) != hipSuccess) {
        qjx_batch_destroy(b);
        return fail(QJX_ERR_HIP, "stream creation failed");
    }
    int rc = qjx::runner_create(ctxs[0]->prec, ctxs[0]->gemm_kind, b->stream, (int)total, &b->runner);
    if (rc) {
        qjx_batch_destroy(b);
        return rc;
    }
    if (const char *env = std::getenv("QJX_GRAPH"))
        if (std::string(env) == "0") b->runner->graphs = false;
    *out = b;
    return QJX_OK;
}

static int batch_run(qjx_batch *b, int n, bool periods, int is_obs)
{
    if (!b) return fail(QJX_ERR_ARG, "null batch");
    if (n < 0) return fail(QJX_ERR_ARG, "negative step count");
    if (n == 0) return QJX_OK;
    std::vector<qjx::Ens> table;
    for (qjx_ctx *c : b->ctx) {
        HIP_TRY(hipStreamSynchronize(c->stream));      // the member's own work is done
        const long long steps = periods ? (long long)n * c->S : n;
        if (steps > (1LL << 30)) return fail(QJX_ERR_ARG, "step count too large");
        qjx::launch_seed(c->st, c->M, (int)steps, c->lists[0], c->cnts[0], c->D, b->stream);
        c->desc.is_obs = is_obs;
        table.push_back(c->desc);
    }
    HIP_TRY(hipGetLastError());
    int rc = qjx::runner_set_table(b->runner, table.data(), (int)table.size());
    if (rc) return rc;
    rc = qjx::runner_run(b->runner, nullptr);
    if (rc) return rc;
    for (qjx_ctx *c : b->ctx) {
        rc = check_err(c);
        if (rc) return rc;
        if (is_obs && c->ens.jump > 0) {
            rc = drain_log(c);
            if (rc) return rc;
        }
    }
    return QJX_OK;
}

extern "C" int qjx_batch_run_steps(qjx_batch *b, int n_steps, int is_obs)
{
    return batch_run(b, n_steps, false, is_obs);
}

extern "C" int qjx_batch_run_periods(qjx_batch *b, int n_periods, int is_obs)
{
    return batch_run(b, n_periods, true, is_obs);
}

extern "C" int qjx_batch_get_stats(qjx_batch *b, qjx_stats *out)
{
    if (!b || !out) return fail(QJX_ERR_ARG, "null argument");
    *out = qjx_stats{};
    for (qjx_ctx *c : b->ctx) {
        unsigned long long h[3] = {0, 0, 0};
        HIP_TRY(hipMemcpy(h, c->counters, sizeof h, hipMemcpyDeviceToHost));
        out->matvecs += (int64_t)h[0];
        out->jumps += (int64_t)h[1];
        out->gemm_flops += 8.0 * double(c->d) * double(c->d) * double(h[0]);
        out->gemm_band_flops += double(h[2]) * 128.0;
    }
    const qjx::Runner *r = b->runner;
    out->iterations = r->iterations;
    out->gemm_ms = r->gemm_ms;
    out->decide_ms = r->tail_ms + r->plan_ms;
    out->gemm_launches = r->gemm_launches;
    out->gemm_bn = r->bn;
    out->gemm_grid = r->grid.gemm;
    out->gemm_waves = r->grid.waves;
    return QJX_OK;
}

extern "C" int qjx_batch_set_profiling(qjx_batch *b, int on)
{
    if (!b) return fail(QJX_ERR_ARG, "null batch");
    qjx::Runner *r = b->runner;
    r->profiling = on != 0;
    r->plan_ms = r->gemm_ms = r->tail_ms = 0.0;
    r->gemm_launches = 0;
    return QJX_OK;
}

// ---------------------------------------------------------------------------
// Runner: iterations over a table of ensembles
// ---------------------------------------------------------------------------
namespace qjx {

int runner_create(int prec, int gemm_kind, hipStream_t stream, int capacity, Runner **out)
{
    *out = nullptr;
    Runner *r = new Runner;
    r->prec = prec;
    r->gemm_kind = gemm_kind;
    r->stream = stream;
    r->capacity = capacity;
    auto bail = [&](int code) {
        runner_destroy(r);
        return code;
    };
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return bail(fail(QJX_ERR_HIP, "device query failed"));
    r->cus = cus;
    int bn = gemm_kind == kGemmFma ? 64 : 128;
    r->autotune = gemm_kind != kGemmFma;
    if (const char *env = std::getenv("QJX_TUNE"))
        if (std::string(env) == "0") r->autotune = false;
    if (const char *env = std::getenv("QJX_BN")) {
        const std::string v(env);
        if (gemm_kind != kGemmFma &&
            (v == "64" || v == "128" || (v == "256" && prec != kFP64) || (v == "32" && prec == kFP64))) {
            bn = std::stoi(v);
            r->autotune = false;
        }
    }
    int waves = bn == 256 ? 8 : 4;       // the 256-column tile has only an 8-wave kernel
    if (const char *env = std::getenv("QJX_WAVES")) {
        if (prec == kBF16x3 && gemm_kind == kGemmMfma && std::string(env) == "8") {
            waves = 8;
            bn = bn == 256 ? 256 : 128;
            r->autotune = false;
        }
    }
    runner_apply_shape(r, bn, gemm_blocks_per_cu(prec, gemm_kind, bn, waves), waves);
    // super-columns of two column tiles: on the d = 1001 workload (1e5 trajectories)
    // the FP32 and 3xBF16 GEMMs run 1% faster than with one and than with three or
    // four (the same results: the tile order does not enter any sum)
    r->sc = 2;
    if (const char *env = std::getenv("QJX_SC")) r->sc = std::max(1, std::min(16, std::atoi(env)));
    // complex-pair GEMM: chunks of two row tiles per work unit to start with.  On the
    // d = 1001 workload (1e5 trajectories) FP32 runs 4.8% and 3xBF16 2.7% faster than
    // with one row tile per unit, and chunks of 4 and 8 are slower than 2; small tables
    // run faster with one, so the tuner tries both (the same results: the chunking
    // does not enter any sum)
    r->rc = cp_layout(prec, gemm_kind) ? 2 : 1;
    r->tune_rc = cp_layout(prec, gemm_kind);
    if (const char *env = std::getenv("QJX_RC"))
        if (cp_layout(prec, gemm_kind)) {
            r->rc = std::max(1, std::min(8, std::atoi(env)));
            r->tune_rc = false;
        }
    r->grid.decide = std::max(1, (capacity + 255) / 256);   // one entry per thread
    r->grid.jump = std::max(1, std::min(capacity, 2048));
    if (hipMalloc((void **)&r->d_ens, kMaxEns * sizeof(Ens)) != hipSuccess ||
        hipMalloc((void **)&r->d_nens, sizeof(int)) != hipSuccess ||
        hipMalloc((void **)&r->d_plan, sizeof(Plan)) != hipSuccess ||
        hipMalloc((void **)&r->d_jl, std::max(capacity, 1) * sizeof(JumpEnt)) != hipSuccess ||
        hipMalloc((void **)&r->d_jcnt, sizeof(int)) != hipSuccess ||
        hipHostMalloc((void **)&r->h_total, sizeof(int), hipHostMallocDefault) != hipSuccess)
        return bail(fail(QJX_ERR_NOMEM, "runner buffers do not fit"));
    if (hipMemsetAsync(r->d_plan, 0, sizeof(Plan), stream) != hipSuccess ||
        hipMemsetAsync(r->d_jcnt, 0, sizeof(int), stream) != hipSuccess ||
        hipMemsetAsync(r->d_nens, 0, sizeof(int), stream) != hipSuccess)
        return bail(fail(QJX_ERR_HIP, "runner init failed"));
    *out = r;
    return QJX_OK;
}

void runner_destroy(Runner *r)
{
    if (!r) return;
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    if (r->exec) (void)hipGraphExecDestroy(r->exec);
    if (r->graph) (void)hipGraphDestroy(r->graph);
    for (hipEvent_t e : r->ev) (void)hipEventDestroy(e);
    dev_free(r->d_ens);
    dev_free(r->d_nens);
    dev_free(r->d_plan);
    dev_free(r->d_jl);
    dev_free(r->d_jcnt);
    if (r->h_total) (void)hipHostFree(r->h_total);
    delete r;
}

int runner_set_table(Runner *r, const Ens *ens, int n)
{
    if (n < 0 || n > kMaxEns) return fail(QJX_ERR_ARG, "too many ensembles in one runner");
    long long tot = 0;
    for (int i = 0; i < n; i++) tot += ens[i].M;
    if (tot > r->capacity) return fail(QJX_ERR_ARG, "runner capacity exceeded");
    if (n) HIP_TRY(hipMemcpyAsync(r->d_ens, ens, n * sizeof(Ens), hipMemcpyHostToDevice, r->stream));
    r->nens = n;
    r->h_ens.assign(ens, ens + n);
    // the jump kernel's LDS follows the table's largest dimension; a larger one
    // needs a new capture of the iteration graph
    int jdim = 64;
    for (int i = 0; i < n; i++) jdim = std::max(jdim, ens[i].d);
    if (jdim > r->grid.jdim || (r->grid.jdim == kMaxDim && jdim < kMaxDim)) {
        r->grid.jdim = jdim;
        if (r->exec) {
            HIP_TRY(hipStreamSynchronize(r->stream));
            (void)hipGraphExecDestroy(r->exec);
            r->exec = nullptr;
        }
        if (r->graph) {
            (void)hipGraphDestroy(r->graph);
            r->graph = nullptr;
        }
    }
    // the count goes through pinned memory so the copy is stream-ordered
    HIP_TRY(hipStreamSynchronize(r->stream));
    *r->h_total = n;
    HIP_TRY(hipMemcpyAsync(r->d_nens, r->h_total, sizeof(int), hipMemcpyHostToDevice, r->stream));
    HIP_TRY(hipStreamSynchronize(r->stream));
    return QJX_OK;
}

static RunArgs run_args(Runner *r, int parity)
{
    RunArgs a;
    a.ens = r->d_ens;
    a.nens = r->d_nens;
    a.plan = r->d_plan;
    a.jl = r->d_jl;
    a.jcnt = r->d_jcnt;
    a.parity = parity;
    a.bn = r->bn;
    a.sc = r->sc;
    a.rc = r->rc;
    return a;
}

void runner_apply_shape(Runner *r, int bn, int wg_per_cu, int waves, int rc)
{
    int g = r->cus * std::max(1, wg_per_cu);
    g = std::max(8, g / 8 * 8);          // a multiple of the 8 XCDs
    if (rc <= 0) rc = r->rc;
    if (r->bn == bn && r->grid.gemm == g && r->grid.waves == waves && r->rc == rc) return;
    r->rc = rc;
    r->bn = bn;
    r->wg_per_cu = std::max(1, wg_per_cu);
    r->grid.gemm = g;
    r->grid.waves = waves;
    if (r->exec) {
        (void)hipStreamSynchronize(r->stream);
        (void)hipGraphExecDestroy(r->exec);
        r->exec = nullptr;
    }
    if (r->graph) {
        (void)hipGraphDestroy(r->graph);
        r->graph = nullptr;
    }
}

// Process-wide cache of tuned shapes, keyed by precision, the largest operand
// dimension of the table and log2 of its trajectory count.
struct ShapeKey {
    int prec, gemm_kind, lmax, log2m;
    bool operator<(const ShapeKey &o) const
    {
        return std::tie(prec, gemm_kind, lmax, log2m) < std::tie(o.prec, o.gemm_kind, o.lmax, o.log2m);
    }
};
static std::mutex g_shape_mu;
struct Shape {
    int bn, wpc, waves, rc;
};
static std::map<ShapeKey, Shape> g_shapes;

static ShapeKey shape_key(const Runner *r)
{
    int lmax = 0;
    long long m = 0;
    for (const Ens &E : r->h_ens) {
        lmax = std::max(lmax, E.L);
        m += E.M;
    }
    int lg = 0;
    while ((1LL << (lg + 1)) <= m) lg++;
    return ShapeKey{r->prec, r->gemm_kind, lmax, lg};
}

static int read_cols(Runner *r, unsigned long long *cols)
{
    HIP_TRY(hipMemcpyAsync(cols, &r->d_plan->cols, sizeof *cols, hipMemcpyDeviceToHost, r->stream));
    HIP_TRY(hipStreamSynchronize(r->stream));
    return QJX_OK;
}

// Matrix-core work the table's GEMMs have executed so far (units of 128 flops)
static int read_work(Runner *r, unsigned long long *work)
{
    unsigned long long sum = 0, v = 0;
    for (const Ens &E : r->h_ens) {
        HIP_TRY(hipMemcpyAsync(&v, E.counters + 2, sizeof v, hipMemcpyDeviceToHost, r->stream));
        HIP_TRY(hipStreamSynchronize(r->stream));
        sum += v;
    }
    *work = sum;
    return QJX_OK;
}

// Optional events recorded around the launches of one iteration.
struct IterEvents {
    hipEvent_t before_plan = nullptr, before_gemm = nullptr, after_gemm = nullptr, after_tail = nullptr;
};

// One iteration (plan, GEMM, decide + jump) with each launch checked on its own: a
// launch the runtime rejects is reported only through hipGetLastError, and inside a
// stream capture it would otherwise surface as an invalidated capture with no cause.
// On failure `what` names the step.
static hipError_t iteration_checked(Runner *r, const RunArgs &a, std::string *what, const IterEvents &ev = {})
{
    hipError_t e = hipSuccess;
    auto record = [&](hipEvent_t x) {
        if (x && (e = hipEventRecord(x, r->stream)) != hipSuccess) *what = "event record";
        return e == hipSuccess;
    };
    auto launched = [&](const char *name) {
        if ((e = hipGetLastError()) != hipSuccess) *what = name;
        return e == hipSuccess;
    };
    if (!record(ev.before_plan)) return e;
    launch_plan(a, r->stream);
    if (!launched("plan launch") || !record(ev.before_gemm)) return e;
    launch_gemm(r->prec, r->gemm_kind, a, r->grid, r->stream);
    if (!launched("GEMM launch") || !record(ev.after_gemm)) return e;
    launch_tail(r->prec, a, r->grid, r->stream);
    if (!launched("decide/jump launch") || !record(ev.after_tail)) return e;
    return hipSuccess;
}

static int iteration(Runner *r, const RunArgs &a, const IterEvents &ev = {})
{
    std::string what;
    const hipError_t e = iteration_checked(r, a, &what, ev);
    return e == hipSuccess ? QJX_OK : fail(QJX_ERR_HIP, what + ": " + hipGetErrorString(e));
}

// Captures two iterations (parity 0 then 1) into the runner's graph; every argument
// lives in device memory, so later table changes need no new capture.  A HIP call
// on another thread that synchronises with this stream invalidates the capture,
// and the runtime can leave the stream in that state, so an invalidated capture is
// an error: owners with other threads defer captures while those threads work.
static int capture_iterations(Runner *r)
{
    HIP_TRY(hipGetLastError());            // an earlier launch error is reported as itself
    HIP_TRY(hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal));
    std::string what;
    hipError_t le = iteration_checked(r, run_args(r, 0), &what);
    if (le == hipSuccess) le = iteration_checked(r, run_args(r, 1), &what);
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(r->stream, &g);   // always ends the capture
    if (le == hipSuccess && ee == hipSuccess) {
        const hipError_t ie = hipGraphInstantiate(&r->exec, g, nullptr, nullptr, 0);
        if (ie != hipSuccess) {
            (void)hipGraphDestroy(g);
            r->exec = nullptr;
            return fail(ie == hipErrorOutOfMemory ? QJX_ERR_NOMEM : QJX_ERR_HIP,
                        std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
        }
        r->graph = g;
        return QJX_OK;
    }
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    const bool invalidated = (le == hipSuccess || le == hipErrorStreamCaptureInvalidated) &&
                             (ee == hipSuccess || ee == hipErrorStreamCaptureInvalidated);
    if (invalidated)
        return fail(QJX_ERR_HIP, "capture of the iteration graph was invalidated: a HIP call on another "
                                 "thread synchronised with the stream during the capture (keep such calls "
                                 "out of the capture, or set QJX_GRAPH=0)");
    return fail(QJX_ERR_HIP, "capture of the iteration graph failed: " +
                                 (le != hipSuccess ? what + ": " + hipGetErrorString(le)
                                                   : std::string(hipGetErrorString(ee))));
}

// `iters` iterations with the GEMM bracketed by events: device ms of the GEMM
// launches and list entries processed.
static int timed_iterations(Runner *r, int iters, double *ms, unsigned long long *cols, unsigned long long *work)
{
    if (r->ev.size() < 2 * (size_t)iters) {
        const size_t old = r->ev.size();
        r->ev.resize(2 * (size_t)iters);
        for (size_t i = old; i < r->ev.size(); i++) HIP_TRY(hipEventCreate(&r->ev[i]));
    }
    unsigned long long c0 = 0, c1 = 0, w0 = 0, w1 = 0;
    int rc = read_cols(r, &c0);
    if (rc) return rc;
    if ((rc = read_work(r, &w0))) return rc;
    for (int k = 0; k < iters; k++) {
        IterEvents ev;
        ev.before_gemm = r->ev[2 * k];
        ev.after_gemm = r->ev[2 * k + 1];
        if ((rc = iteration(r, run_args(r, k & 1), ev))) return rc;
    }
    rc = read_cols(r, &c1);
    if (rc) return rc;
    if ((rc = read_work(r, &w1))) return rc;
    double t = 0.0;
    for (int k = 0; k < iters; k++) {
        float g = 0.f;
        HIP_TRY(hipEventElapsedTime(&g, r->ev[2 * k], r->ev[2 * k + 1]));
        t += g;
    }
    *ms = t;
    *cols = c1 - c0;
    *work = w1 - w0;
    r->iterations += iters;
    if (r->profiling) {
        r->gemm_ms += t;
        r->gemm_launches += iters;
    }
    return QJX_OK;
}

// A shape this process has already tuned for the runner's table shape and form:
// applied, and the runner does not tune.
static bool apply_cached_shape(Runner *r)
{
    if (!r->autotune || r->tuned || r->nens == 0) return false;
    const ShapeKey key = shape_key(r);
    std::lock_guard<std::mutex> lk(g_shape_mu);
    auto it = g_shapes.find(key);
    if (it == g_shapes.end()) return false;
    runner_apply_shape(r, it->second.bn, it->second.wpc, it->second.waves, it->second.rc);
    r->tuned = true;
    r->cached = true;
    return true;
}

int runner_autotune(Runner *r, bool again)
{
    if (!r->autotune || r->nens == 0) return QJX_OK;
    if (r->tuned && (!again || r->cached)) return QJX_OK;
    if (apply_cached_shape(r)) return QJX_OK;
    const ShapeKey key = shape_key(r);
    struct Cand {
        int bn, wpc, waves, rc;
        double ms;
        unsigned long long cols, work;
    };
    std::vector<Cand> cand;
    // row tiles per complex-pair work unit: two suit large tables (d = 1001, 1e5
    // trajectories: 7% faster than one), one suits small ones, whose few units leave
    // most workgroups idle (d = 101 to 501 with 1000 trajectories: 26-38% faster)
    std::vector<int> rcs{r->rc};
    if (r->tune_rc) rcs = {2, 1};
    auto add = [&](int bn, int wpc, int waves) {
        for (int rc : rcs) cand.push_back({bn, wpc, waves, rc, 0.0, 0, 0});
    };
    // FP64: also 32-column tiles, whose smaller tiles give small tables more workgroups
    std::vector<int> bns{128, 64};
    if (r->prec == kFP64) bns.push_back(32);
    for (int bn : bns) {
        const int occ = gemm_blocks_per_cu(r->prec, r->gemm_kind, bn, 4);
        add(bn, occ, 4);
        if (occ > 1) add(bn, occ / 2, 4);
    }
    if (r->prec == kBF16x3) {
        const int occ = gemm_blocks_per_cu(r->prec, r->gemm_kind, 128, 8);
        add(128, occ, 8);
    }
    if (r->prec != kFP64) {                // 256-column tiles (8 waves) for the FP32 forms
        const int occ = gemm_blocks_per_cu(r->prec, r->gemm_kind, 256, 8);
        add(256, occ, 8);
    }
    const int saved_bn = r->bn, saved_wpc = r->wg_per_cu, saved_waves = r->grid.waves, saved_rc = r->rc;
    constexpr int kIters = 4, kRounds = 3;
    for (int round = 0; round < kRounds; round++) {
        for (Cand &c : cand) {
            runner_apply_shape(r, c.bn, c.wpc, c.waves, c.rc);
            double ms = 0.0;
            unsigned long long cols = 0, work = 0;
            int rc = timed_iterations(r, kIters, &ms, &cols, &work);
            if (rc) return rc;
            c.ms += ms;
            c.cols += cols;
            c.work += work;
            int total = 0;
            rc = runner_next_total(r, &total);
            if (rc) return rc;
            if (total == 0) {      // the table ran out of work: try again next time
                runner_apply_shape(r, saved_bn, saved_wpc, saved_waves, saved_rc);
                return QJX_OK;
            }
        }
    }
    // GEMM time per executed matrix-core flop: a column of a wide-band group costs
    // several of a deep level's, and each candidate's trial iterations hold a
    // different mix of groups, so time per column would favour whichever candidate
    // drew the narrow ones (time per column when no work was counted)
    auto cost = [](const Cand &c) { return c.work > 0 ? c.ms / (double)c.work : c.ms / (double)c.cols; };
    const Cand *best = nullptr;
    for (const Cand &c : cand)
        if (c.cols > 0 && (!best || cost(c) < cost(*best))) best = &c;
    if (std::getenv("QJX_TUNE_VERBOSE")) {
        for (const Cand &c : cand)
            std::fprintf(stderr, "qjx tune: bn %d wpc %d waves %d rc %d: %.3f ms, %llu columns, %.4g ms per Gflop%s\n",
                         c.bn, c.wpc, c.waves, c.rc, c.ms, c.cols, c.work > 0 ? c.ms / (c.work * 128.0 * 1e-9) : 0.0,
                         &c == best ? "  <- picked" : "");
    }
    if (!best) {
        runner_apply_shape(r, saved_bn, saved_wpc, saved_waves, saved_rc);
        return QJX_OK;
    }
    runner_apply_shape(r, best->bn, best->wpc, best->waves, best->rc);
    r->tuned = true;
    r->tunes++;
    std::lock_guard<std::mutex> lk(g_shape_mu);
    g_shapes[key] = {best->bn, best->wpc, best->waves, best->rc};
    return QJX_OK;
}

int runner_chunk(Runner *r, int iters)
{
    if (iters <= 0 || (iters & 1)) return fail(QJX_ERR_ARG, "chunk length must be even");
    if (r->profiling) {
        if (r->ev.size() < (size_t)4 * iters) {
            const size_t old = r->ev.size();
            r->ev.resize(4 * (size_t)iters);
            for (size_t i = old; i < r->ev.size(); i++) HIP_TRY(hipEventCreate(&r->ev[i]));
        }
        for (int k = 0; k < iters; k++) {
            IterEvents ev;
            ev.before_plan = r->ev[4 * k];
            ev.before_gemm = r->ev[4 * k + 1];
            ev.after_gemm = r->ev[4 * k + 2];
            ev.after_tail = r->ev[4 * k + 3];
            const int rc = iteration(r, run_args(r, k & 1), ev);
            if (rc) return rc;
        }
        HIP_TRY(hipStreamSynchronize(r->stream));
        for (int k = 0; k < iters; k++) {
            float p = 0.f, g = 0.f, t = 0.f;
            HIP_TRY(hipEventElapsedTime(&p, r->ev[4 * k], r->ev[4 * k + 1]));
            HIP_TRY(hipEventElapsedTime(&g, r->ev[4 * k + 1], r->ev[4 * k + 2]));
            HIP_TRY(hipEventElapsedTime(&t, r->ev[4 * k + 2], r->ev[4 * k + 3]));
            r->plan_ms += p;
            r->gemm_ms += g;
            r->tail_ms += t;
        }
        r->gemm_launches += iters;
    } else {
        if (r->graphs && !r->exec && !r->defer_capture) {
            const int rc = capture_iterations(r);
            if (rc) return rc;
        }
        if (r->exec) {
            for (int k = 0; k < iters; k += 2) HIP_TRY(hipGraphLaunch(r->exec, r->stream));
        } else {
            for (int k = 0; k < iters; k++) {
                const int rc = iteration(r, run_args(r, k & 1));
                if (rc) return rc;
            }
        }
    }
    r->iterations += iters;
    return QJX_OK;
}

int runner_next_total(Runner *r, int *total)
{
    HIP_TRY(hipMemcpyAsync(r->h_total, &r->d_plan->next_total[0], sizeof(int), hipMemcpyDeviceToHost, r->stream));
    HIP_TRY(hipStreamSynchronize(r->stream));
    *total = *r->h_total;
    return QJX_OK;
}

// Iterations a runner runs before it tunes its launch shape (chunks of 8, 16, ...,
// so tuning starts at the first chunk boundary past this count), and measures
// again at 8 and 64 times that count, when the bisection has spread the table over
// its groups.  On the d = 1001 workload (1e5 trajectories) every trajectory is
// still at level 0 after 512 iterations from the start state: there the 128- and
// 256-column FP32 tiles tie, while over a period of the run's later mix the
// 256-column tiles are 3.5% faster.  QJX_TUNE_AFTER overrides the count.
static long long tune_after()
{
    static const long long n = [] {
        const char *env = std::getenv("QJX_TUNE_AFTER");
        return env ? std::max(0LL, std::atoll(env)) : 512LL;
    }();
    return n;
}

int runner_run(Runner *r, long long *iterations)
{
    int chunk = 8;
    long long done = 0;
    // a table shape tuned before in this process runs its shape from the first
    // iteration on
    (void)apply_cached_shape(r);
    for (;;) {
        int rc = runner_chunk(r, chunk);
        if (rc) return rc;
        done += chunk;
        int total = 0;
        rc = runner_next_total(r, &total);
        if (rc) return rc;
        if (total == 0) break;
        long long traj = 0;
        for (const Ens &E : r->h_ens) traj += E.M;
        if (r->autotune && !r->cached && r->tunes < 3 && r->iterations >= tune_after() << (3 * r->tunes) &&
            2LL * total >= traj) {
            // shape the GEMM, and shape it again as the bisection spreads the table
            // over its groups: the first iterations from a fresh start are level-0
            // steps of the whole table (wide bands), where other shapes win than in
            // the run's later mix of deep levels.  Never in the tail of a call, when
            // most trajectories have reached their target: the few left are not the
            // run's mix
            const long long before = r->iterations;
            rc = runner_autotune(r, r->tunes > 0);
            if (rc) return rc;
            done += r->iterations - before;
            rc = runner_next_total(r, &total);
            if (rc) return rc;
            if (total == 0) break;
        }
        chunk = std::min(chunk * 2, 256);
    }
    if (iterations) *iterations = done;
    return QJX_OK;
}

}  // namespace qjx

// ---------------------------------------------------------------------------
// Helpers for the job pool
// ---------------------------------------------------------------------------
namespace qjx {

static int padded_dim(int N) { return (N + 1 + 63) / 64 * 64; }

size_t model_footprint(const qjx_model_desc *m, int gemm_kind)
{
    const size_t L = 2 * (size_t)padded_dim(m->N);
    const size_t rows = cp_layout(m->precision, gemm_kind) ? L / 2 : L;
    const size_t pe = m->precision == QJX_FP64 ? 8 : 4;     // BF16x3: hi + lo = 4 bytes
    const size_t d = m->N + 1;
    // propagator tree + the build workspace (9 dense complex d x d, freed after the build)
    return (size_t)2 * m->ex_deep * rows * L * pe + 9 * d * d * 16 + (1 << 20);
}

size_t ensemble_footprint(const qjx_model_desc *m, int M, int gemm_kind)
{
    const size_t L = 2 * (size_t)padded_dim(m->N);
    const size_t es = m->precision == QJX_FP64 ? 8 : 4;
    const size_t R = L / gemm_bm(gemm_kind);
    const size_t G = 2 * (size_t)m->ex_deep;
    return (size_t)M * (2 * L * es + 2 * G * 4 + R * 8 + sizeof(TrajState) + 7 * 8) + (1 << 20);
}

int ctx_seed(qjx_ctx *c, int n_steps, int is_obs)
{
    launch_seed(c->st, c->M, n_steps, c->lists[0], c->cnts[0], c->D, c->stream);
    HIP_TRY(hipGetLastError());
    c->desc.is_obs = is_obs;
    return QJX_OK;
}

}  // namespace qjx
