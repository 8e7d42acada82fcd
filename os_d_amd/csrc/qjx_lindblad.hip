// qjx_lindblad.hip -- the dimer's Lindblad equation integrated in rho space on the GPU
// (FP64), the ensemble-level counterpart of the trajectory engine.
//
//   d rho / dt = -i [H(t), rho] + g (L rho L^dag - 1/2 {L^dag L, rho}),  g = gamma / N,
//   H(t) = H + E(t) H_drv,  E(t) = E + A on the first half period, E - A on the second.
//
// The reference integrates this equation with RK4 in the SU(N) f-basis
// (CQdiss_os_dimer/CQdiss/calcODE.cpp:310-389 calcODE_std; rho recovered by
// CalcRho.cpp:4-67).  A fixed linear change of variables commutes with RK4, so RK4
// on rho itself is the same scheme.  H and L are tridiagonal and L^dag L
// pentadiagonal, so a right-hand side is a 5-point stencil in each index of rho:
// the integration is HBM-bound, one pass over rho per RK4 stage.
//
// Each stage kernel evaluates the right-hand side with exactly the operations, in
// exactly the order, of the CPU restatement used as the oracle, with floating-point
// contraction off, so both produce the same bits.  The drive is held constant over
// each step and steps are aligned with the half periods (CQdiss evaluates
// sign(sin(omega t)) per stage, taking one stage from the neighbouring half period
// at each switch).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/qjx.h"
#include "qjx_internal.h"
#include "qjx_kernels.h"
#include "qjx_runner.h"

#pragma clang fp contract(off)

namespace qjx {

// band matrices of the model on the device: H diagonal and off-diagonal, H_drv
// diagonal, L (L_kk = ld, L_{k+1,k} = -lo, L_{k,k+1} = lo), L^dag L bands q0..q2
struct LBands {
    const double *hd, *hdrv, *ho, *ld, *lo, *q0, *q1, *q2;
    double g;
    int d;
};

namespace {

typedef double2 c2;

__device__ inline c2 cadd(c2 a, c2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline c2 csub(c2 a, c2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline c2 rmul(double s, c2 a) { return make_double2(s * a.x, s * a.y); }
__device__ inline c2 mulr(c2 a, double s) { return make_double2(a.x * s, a.y * s); }

// (L x)_i over the rows of column j (L_ii = ld, L_{i,i-1} = -lo[i-1], L_{i,i+1} = lo[i])
__device__ inline c2 lrow(const LBands &B, const c2 *x, int i, int j)
{
    const int d = B.d;
    c2 s = rmul(B.ld[i], x[(size_t)i * d + j]);
    if (i > 0) s = cadd(s, rmul(-B.lo[i - 1], x[(size_t)(i - 1) * d + j]));
    if (i + 1 < d) s = cadd(s, rmul(B.lo[i], x[(size_t)(i + 1) * d + j]));
    return s;
}

// right-hand side at (i, j); terms in the order of the CPU restatement
__device__ inline c2 rhs_at(const LBands &B, double E, const c2 *x, int i, int j)
{
    const int d = B.d;
    const size_t ij = (size_t)i * d + j;
    // (L x L^T)_ij from T = L x: T_ij, T_i,j-1, T_i,j+1
    c2 lrl = rmul(B.ld[j], lrow(B, x, i, j));
    if (j > 0) lrl = cadd(lrl, rmul(-B.lo[j - 1], lrow(B, x, i, j - 1)));
    if (j + 1 < d) lrl = cadd(lrl, rmul(B.lo[j], lrow(B, x, i, j + 1)));
    const double hdi = B.hd[i] + E * B.hdrv[i], hdj = B.hd[j] + E * B.hdrv[j];
    c2 hr = rmul(hdi, x[ij]);
    if (i > 0) hr = cadd(hr, rmul(B.ho[i - 1], x[ij - d]));
    if (i + 1 < d) hr = cadd(hr, rmul(B.ho[i], x[ij + d]));
    c2 rh = mulr(x[ij], hdj);
    if (j > 0) rh = cadd(rh, mulr(x[ij - 1], B.ho[j - 1]));
    if (j + 1 < d) rh = cadd(rh, mulr(x[ij + 1], B.ho[j]));
    c2 qr = rmul(B.q0[i], x[ij]);
    if (i + 1 < d) qr = cadd(qr, rmul(B.q1[i], x[ij + d]));
    if (i + 2 < d) qr = cadd(qr, rmul(B.q2[i], x[ij + 2 * (size_t)d]));
    if (i >= 1) qr = cadd(qr, rmul(B.q1[i - 1], x[ij - d]));
    if (i >= 2) qr = cadd(qr, rmul(B.q2[i - 2], x[ij - 2 * (size_t)d]));
    c2 rq = mulr(x[ij], B.q0[j]);
    if (j + 1 < d) rq = cadd(rq, mulr(x[ij + 1], B.q1[j]));
    if (j + 2 < d) rq = cadd(rq, mulr(x[ij + 2], B.q2[j]));
    if (j >= 1) rq = cadd(rq, mulr(x[ij - 1], B.q1[j - 1]));
    if (j >= 2) rq = cadd(rq, mulr(x[ij - 2], B.q2[j - 2]));
    // (0 - i) * (hr - rh): the complex product with exact zeros, as std::complex
    const c2 c = csub(hr, rh);
    const c2 mi = make_double2(0.0 * c.x - (-1.0) * c.y, 0.0 * c.y + (-1.0) * c.x);
    const c2 diss = rmul(B.g, csub(lrl, rmul(0.5, cadd(qr, rq))));
    return cadd(mi, diss);
}

// stage s of one RK4 step: k = f(x); s < 3: v = rho + c k (k kept for the final
// combination); s = 3: rho += h6 (((k1 + 2 k2) + 2 k3) + k)
__global__ __launch_bounds__(256) void rk4_stage_kernel(LBands B, double E, const c2 *x, c2 *rho, c2 *kout,
                                                        c2 *vout, double c, const c2 *k1, const c2 *k2,
                                                        const c2 *k3, double h6, int last)
{
    const size_t n = (size_t)B.d * B.d;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(e / B.d), j = (int)(e % B.d);
        const c2 k = rhs_at(B, E, x, i, j);
        if (!last) {
            kout[e] = k;
            vout[e] = cadd(rho[e], rmul(c, k));
        } else {
            const c2 sum = cadd(cadd(cadd(k1[e], rmul(2.0, k2[e])), rmul(2.0, k3[e])), k);
            rho[e] = cadd(rho[e], rmul(h6, sum));
        }
    }
}

// <k>, <(k - <k>)^2>, <H + E0 H_drv>, adr_k = rho_kk (fixed-order sums, one workgroup)
__global__ __launch_bounds__(256) void lb_obs_kernel(LBands B, double E0, const c2 *rho, double *out, double *adr)
{
    __shared__ double s[3][256];
    const int d = B.d, t = threadIdx.x;
    double mean = 0.0, en = 0.0;
    for (int k = t; k < d; k += 256) {
        const double p = rho[(size_t)k * d + k].x;
        if (adr) adr[k] = p;
        mean += (double)k * p;
        en += (B.hd[k] + E0 * B.hdrv[k]) * p;
        if (k + 1 < d) en += B.ho[k] * (rho[(size_t)k * d + k + 1].x + rho[(size_t)(k + 1) * d + k].x);
    }
    s[0][t] = mean;
    s[1][t] = en;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off) {
            s[0][t] += s[0][t + off];
            s[1][t] += s[1][t + off];
        }
        __syncthreads();
    }
    const double m = s[0][0];
    __syncthreads();
    double m2 = 0.0;
    for (int k = t; k < d; k += 256) {
        const double p = rho[(size_t)k * d + k].x;
        m2 += ((double)k - m) * ((double)k - m) * p;
    }
    s[2][t] = m2;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off) s[2][t] += s[2][t + off];
        __syncthreads();
    }
    if (t == 0) {
        out[0] = s[0][0];
        out[1] = s[2][0];
        out[2] = s[1][0];
    }
}

}  // namespace
}  // namespace qjx

using qjx::set_last_error;

struct qjx_lindblad {
    int d = 0, steps = 0;
    double T = 0.0, h = 0.0, E0 = 0.0, E1 = 0.0, Estat = 0.0;
    hipStream_t stream = nullptr;
    double *bands = nullptr;          // hd, hdrv, ho, ld, lo, q0, q1, q2 (d each)
    double2 *buf = nullptr;           // rho, va, vb, k1, k2, k3 (d x d each)
    double *obs = nullptr;            // 3 observables + purity + adr[d]
    qjx::LBands B{};
    hipGraphExec_t exec[2] = {nullptr, nullptr};   // one RK4 step with E0 / E1
    hipGraph_t graph[2] = {nullptr, nullptr};
};

namespace {

int lb_fail(int code, const std::string &msg)
{
    set_last_error(msg);
    return code;
}

#define LB_TRY(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return lb_fail(e_ == hipErrorOutOfMemory ? QJX_ERR_NOMEM : QJX_ERR_HIP,           \
                           std::string(#expr) + ": " + hipGetErrorString(e_));                \
    } while (0)

size_t nsq(const qjx_lindblad *l) { return (size_t)l->d * l->d; }
double2 *rho_of(qjx_lindblad *l) { return l->buf; }

// the band matrices of H, H_drv, L and L^dag L (reference newdel.cpp:450-498,
// 1454-1498), summed as the CPU restatement sums them
void build_bands(const qjx_model_desc *m, int d, std::vector<double> &h)
{
    h.assign((size_t)8 * d, 0.0);
    double *hd = h.data(), *hdrv = hd + d, *ho = hd + 2 * d, *ld = hd + 3 * d, *lo = hd + 4 * d;
    double *q0 = hd + 5 * d, *q1 = hd + 6 * d, *q2 = hd + 7 * d;
    const double U = m->prm_U / double(d - 1);
    for (int k = 0; k < d; k++) {
        hd[k] = 2.0 * U * double(k * (k - 1) + (d - (k + 1)) * (d - (k + 1) - 1));
        hdrv[k] = double((d - (k + 1)) - k);
        ld[k] = double(k - (d - (k + 1)));
    }
    for (int k = 0; k + 1 < d; k++) {
        ho[k] = -m->prm_J * std::sqrt(double((d - (k + 1)) * (k + 1)));
        lo[k] = std::sqrt(double((d - (k + 1)) * (k + 1)));
    }
    auto Lel = [&](int i, int j) -> double {
        if (i == j) return ld[i];
        if (i == j + 1) return -lo[j];
        if (j == i + 1) return lo[i];
        return 0.0;
    };
    for (int i = 0; i < d; i++)
        for (int off = 0; off <= 2 && i + off < d; off++) {
            const int j = i + off;
            double s = 0.0;
            for (int k = 0; k < d; k++) s += Lel(k, i) * Lel(k, j);
            (off == 0 ? q0 : off == 1 ? q1 : q2)[i] = s;
        }
}

void launch_step(qjx_lindblad *l, double E, hipStream_t s)
{
    const size_t n = nsq(l);
    double2 *rho = l->buf, *va = rho + n, *vb = va + n, *k1 = vb + n, *k2 = k1 + n, *k3 = k2 + n;
    const int grid = (int)std::min<size_t>((n + 255) / 256, 8192);
    const double h = l->h, c2h = 0.5 * h, h6 = h / 6.0;
    qjx::rk4_stage_kernel<<<grid, 256, 0, s>>>(l->B, E, rho, rho, k1, va, c2h, nullptr, nullptr, nullptr, h6, 0);
    qjx::rk4_stage_kernel<<<grid, 256, 0, s>>>(l->B, E, va, rho, k2, vb, c2h, nullptr, nullptr, nullptr, h6, 0);
    qjx::rk4_stage_kernel<<<grid, 256, 0, s>>>(l->B, E, vb, rho, k3, va, h, nullptr, nullptr, nullptr, h6, 0);
    qjx::rk4_stage_kernel<<<grid, 256, 0, s>>>(l->B, E, va, rho, nullptr, nullptr, 0.0, k1, k2, k3, h6, 1);
}

}  // namespace

extern "C" void qjx_lindblad_destroy(qjx_lindblad *l)
{
    if (!l) return;
    if (l->stream) (void)hipStreamSynchronize(l->stream);
    for (int b = 0; b < 2; b++) {
        if (l->exec[b]) (void)hipGraphExecDestroy(l->exec[b]);
        if (l->graph[b]) (void)hipGraphDestroy(l->graph[b]);
    }
    if (l->bands) (void)hipFree(l->bands);
    if (l->buf) (void)hipFree(l->buf);
    if (l->obs) (void)hipFree(l->obs);
    if (l->stream) (void)hipStreamDestroy(l->stream);
    delete l;
}

extern "C" int qjx_lindblad_create(const qjx_model_desc *m, int steps_per_period, qjx_lindblad **out)
{
    if (!m || !out) return lb_fail(QJX_ERR_ARG, "null argument");
    *out = nullptr;
    if (m->N < 1 || m->N + 1 > 4096) return lb_fail(QJX_ERR_ARG, "N must be in [1, 4095]");
    if (steps_per_period < 2 || (steps_per_period % 2)) return lb_fail(QJX_ERR_ARG, "steps_per_period must be even and >= 2");
    if (!(m->drv_freq > 0.0)) return lb_fail(QJX_ERR_ARG, "dimer_drv_freq must be > 0");
    qjx_lindblad *l = new qjx_lindblad;
    l->d = m->N + 1;
    l->steps = steps_per_period;
    l->T = 2.0 * 3.1415926535897932384626433832795 / m->drv_freq;
    l->h = l->T / double(steps_per_period);
    l->E0 = m->prm_E + m->drv_ampl;
    l->E1 = m->prm_E - m->drv_ampl;
    l->Estat = m->prm_E;
    auto bail = [&](int code) {
        qjx_lindblad_destroy(l);
        return code;
    };
    std::vector<double> hb;
    build_bands(m, l->d, hb);
    const int d = l->d;
    const size_t n = nsq(l);
    if (hipStreamCreate(&l->stream) != hipSuccess) return bail(lb_fail(QJX_ERR_HIP, "hipStreamCreate failed"));
    if (hipMalloc((void **)&l->bands, hb.size() * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&l->buf, 6 * n * sizeof(double2)) != hipSuccess ||
        hipMalloc((void **)&l->obs, ((size_t)4 + d) * sizeof(double)) != hipSuccess)
        return bail(lb_fail(QJX_ERR_NOMEM, "Lindblad buffers do not fit in device memory"));
    if (hipMemcpy(l->bands, hb.data(), hb.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return bail(lb_fail(QJX_ERR_HIP, "band upload failed"));
    double *bp = l->bands;
    l->B = qjx::LBands{bp, bp + d, bp + 2 * d, bp + 3 * d, bp + 4 * d, bp + 5 * d, bp + 6 * d, bp + 7 * d,
                       m->diss_gamma / double(d - 1), d};
    // start in the pure state e_0 (start_type 0, start_state 0)
    std::vector<double2> r0(n, make_double2(0.0, 0.0));
    r0[0] = make_double2(1.0, 0.0);
    if (hipMemcpy(l->buf, r0.data(), n * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess)
        return bail(lb_fail(QJX_ERR_HIP, "rho upload failed"));
    *out = l;
    return QJX_OK;
}

extern "C" int qjx_lindblad_set_rho(qjx_lindblad *l, const double *rho)
{
    if (!l || !rho) return lb_fail(QJX_ERR_ARG, "null argument");
    LB_TRY(hipMemcpy(rho_of(l), rho, nsq(l) * sizeof(double2), hipMemcpyHostToDevice));
    return QJX_OK;
}

extern "C" int qjx_lindblad_get_rho(qjx_lindblad *l, double *rho)
{
    if (!l || !rho) return lb_fail(QJX_ERR_ARG, "null argument");
    LB_TRY(hipStreamSynchronize(l->stream));
    LB_TRY(hipMemcpy(rho, rho_of(l), nsq(l) * sizeof(double2), hipMemcpyDeviceToHost));
    return QJX_OK;
}

extern "C" int qjx_lindblad_run_periods(qjx_lindblad *l, int n_periods)
{
    if (!l) return lb_fail(QJX_ERR_ARG, "null argument");
    if (n_periods < 0) return lb_fail(QJX_ERR_ARG, "negative period count");
    if (n_periods == 0) return QJX_OK;
    // one captured RK4 step per half period's drive value, replayed; each launch of
    // the capture is checked, so a rejected launch is reported as itself
    for (int b = 0; b < 2; b++) {
        if (l->exec[b]) continue;
        LB_TRY(hipGetLastError());                 // an earlier launch error, as itself
        LB_TRY(hipStreamBeginCapture(l->stream, hipStreamCaptureModeThreadLocal));
        launch_step(l, b == 0 ? l->E0 : l->E1, l->stream);
        const hipError_t le = hipGetLastError();
        hipGraph_t g = nullptr;
        const hipError_t ee = hipStreamEndCapture(l->stream, &g);   // always ends the capture
        if (le == hipSuccess && ee == hipSuccess) {
            const hipError_t ie = hipGraphInstantiate(&l->exec[b], g, nullptr, nullptr, 0);
            if (ie != hipSuccess) {
                (void)hipGraphDestroy(g);
                l->exec[b] = nullptr;
                return lb_fail(ie == hipErrorOutOfMemory ? QJX_ERR_NOMEM : QJX_ERR_HIP,
                               std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
            }
            l->graph[b] = g;
            continue;
        }
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return lb_fail(QJX_ERR_HIP, std::string("capture of the RK4 step failed: ") +
                                        hipGetErrorString(le != hipSuccess ? le : ee));
    }
    const int half = l->steps / 2;
    for (int p = 0; p < n_periods; p++)
        for (int b = 0; b < 2; b++)
            for (int s = 0; s < half; s++) LB_TRY(hipGraphLaunch(l->exec[b], l->stream));
    LB_TRY(hipStreamSynchronize(l->stream));
    return QJX_OK;
}

extern "C" int qjx_lindblad_observables(qjx_lindblad *l, double *mean, double *m2, double *energy,
                                        double *purity, double *adr)
{
    if (!l) return lb_fail(QJX_ERR_ARG, "null argument");
    const int d = l->d;
    qjx::lb_obs_kernel<<<1, 256, 0, l->stream>>>(l->B, l->Estat, rho_of(l), l->obs, l->obs + 4);
    LB_TRY(hipGetLastError());
    qjx::launch_purity(reinterpret_cast<const double *>(rho_of(l)), (size_t)d * d, l->obs + 3, l->stream);
    LB_TRY(hipGetLastError());
    std::vector<double> h((size_t)4 + d);
    LB_TRY(hipMemcpyAsync(h.data(), l->obs, h.size() * sizeof(double), hipMemcpyDeviceToHost, l->stream));
    LB_TRY(hipStreamSynchronize(l->stream));
    if (mean) *mean = h[0];
    if (m2) *m2 = h[1];
    if (energy) *energy = h[2];
    if (purity) *purity = h[3];
    if (adr)
        for (int k = 0; k < d; k++) adr[k] = h[4 + k];
    return QJX_OK;
}

extern "C" int qjx_lindblad_dim(const qjx_lindblad *l) { return l ? l->d : -1; }
