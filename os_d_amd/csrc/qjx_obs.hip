// qjx_obs.hip -- random observables and the LPN task's per-period bookkeeping on the
// device, in the reference's operation order so that the results equal the CPU
// restatement's bit for bit (FP64 states).
//
// Reference (source/cpp/QJX/QJX/): init_random_obs (newdel.cpp:3652-3722),
// get_random_obs (experiment.cpp:2023-2078), calc_chars_std(_start) /
// calc_chars_lpn(_start) / calc_delta_s / calc_delta_f (core.cpp:665-915),
// copy_trajectory_lpn / var_trajectory_lpn (experiment.cpp:2221-2393),
// lambda_lpn / lambda_lpn_now (experiment.cpp:3065-3134).
//
// Sums the reference leaves to a library (the random observable's zgemv, X X^dag's
// zgemm) run in ascending index order; every multiply and add is rounded on its own
// (no contraction), as in the restatement.  Reductions that the reference performs
// in one sequential loop run on one lane.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/qjx.h"
#include "qjx_ctx.h"
#include "qjx_obs.h"

#pragma clang fp contract(off)

namespace qjx {

// ---------------------------------------------------------------------------
// 10^y and ln x, evaluated identically on the host and the device (libm's pow and
// log may differ between the two in the last place; these stay within 1 ulp of
// the host C library).  10^y = 2^n e^r: t = y ln 10 as a double-double, n the
// nearest integer to t / ln 2, r = t - n ln 2 with ln 2 split into 32 + 53
// significant bits (Cody & Waite), e^r by its Taylor series to degree 14 on
// |r| <= ln(2)/2.  ln x = e ln 2 + 2 atanh s for x = 2^e m, m in [sqrt(1/2),
// sqrt(2)), s = (m - 1)/(m + 1), the series to s^23 arranged as in fdlibm's log
// so that the leading term m - 1 is exact.
// ---------------------------------------------------------------------------
__host__ __device__ inline double lpn_exp10(double y)
{
    const double th = y * 0x1.26bb1bbb55516p+1;
    const double tl = fma(y, 0x1.26bb1bbb55516p+1, -th) + y * -0x1.f48ad494ea3e9p-53;
    const double n = rint(th * 0x1.71547652b82fep+0);
    double r = th - n * 0x1.62e42fee00000p-1;
    r = r - n * 0x1.a39ef35793c76p-33;
    r = r + tl;
    double p = 0x1.93974a8c07c9dp-37;
    p = fma(p, r, 0x1.6124613a86d09p-33);
    p = fma(p, r, 0x1.1eed8eff8d898p-29);
    p = fma(p, r, 0x1.ae64567f544e4p-26);
    p = fma(p, r, 0x1.27e4fb7789f5cp-22);
    p = fma(p, r, 0x1.71de3a556c734p-19);
    p = fma(p, r, 0x1.a01a01a01a01ap-16);
    p = fma(p, r, 0x1.a01a01a01a01ap-13);
    p = fma(p, r, 0x1.6c16c16c16c17p-10);
    p = fma(p, r, 0x1.1111111111111p-7);
    p = fma(p, r, 0x1.5555555555555p-5);
    p = fma(p, r, 0x1.5555555555555p-3);
    p = fma(p, r, 0x1.0000000000000p-1);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)n);
}

__host__ __device__ inline double lpn_log(double x)
{
    int e = 0;
    double m = frexp(x, &e);
    if (m < 0x1.6a09e667f3bcdp-1) {
        m = m * 2.0;
        e -= 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double R = 0x1.642c8590b2164p-4;
    R = fma(R, z, 0x1.8618618618618p-4);
    R = fma(R, z, 0x1.af286bca1af28p-4);
    R = fma(R, z, 0x1.e1e1e1e1e1e1ep-4);
    R = fma(R, z, 0x1.1111111111111p-3);
    R = fma(R, z, 0x1.3b13b13b13b14p-3);
    R = fma(R, z, 0x1.745d1745d1746p-3);
    R = fma(R, z, 0x1.c71c71c71c71cp-3);
    R = fma(R, z, 0x1.2492492492492p-2);
    R = fma(R, z, 0x1.999999999999ap-2);
    R = fma(R, z, 0x1.5555555555555p-1);
    R = R * z;
    const double hfsq = 0.5 * f * f;
    const double de = double(e);
    return de * 0x1.62e42fee00000p-1 - ((hfsq - (s * (hfsq + R) + de * 0x1.a39ef35793c76p-33)) - f);
}

// MKL VSL MCG31m1: x <- a x mod (2^31 - 1); a uniform on [lo, hi) is
// fma(double(x), (hi - lo) (1 / m), lo), after which the stream advances.
constexpr uint64_t kMcg31M = 2147483647ULL;

__host__ __device__ inline double mcg31_next(uint64_t &x, uint64_t stride, double lo, double hi)
{
    const double r = fma(double(x), (hi - lo) * (1.0 / double(kMcg31M)), lo);
    x = x * stride % kMcg31M;
    return r;
}

// ---------------------------------------------------------------------------
// C = A B, ascending k per element, no contraction (64 x 64 tiles, 4 x 4 per thread).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void zgemm_seqk_kernel(int rows, int cols, int K, const double2 *A, int lda,
                                                         const double2 *B, int ldb, double2 *C, int ldc)
{
    __shared__ double2 As[16][65];
    __shared__ double2 Bs[16][65];
    const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
    const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
    double2 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = make_double2(0.0, 0.0);
    for (int k0 = 0; k0 < K; k0 += 16) {
        for (int e = tid; e < 16 * 64; e += 256) {
            const int kk = e / 64, r = e % 64;
            const int k = k0 + kk;
            As[kk][r] = (k < K && r0 + r < rows) ? A[(size_t)(r0 + r) * lda + k] : make_double2(0.0, 0.0);
            Bs[kk][r] = (k < K && c0 + r < cols) ? B[(size_t)k * ldb + c0 + r] : make_double2(0.0, 0.0);
        }
        __syncthreads();
        const int kn = min(16, K - k0);
        for (int kk = 0; kk < kn; kk++) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const double2 a = As[kk][ty + 16 * i];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const double2 b = Bs[kk][tx + 16 * j];
                    acc[i][j].x = acc[i][j].x + (a.x * b.x - a.y * b.y);
                    acc[i][j].y = acc[i][j].y + (a.x * b.y + a.y * b.x);
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int r = r0 + ty + 16 * i, c = c0 + tx + 16 * j;
            if (r < rows && c < cols) C[(size_t)r * ldc + c] = acc[i][j];
        }
}

void launch_zgemm_seqk(int rows, int cols, int K, const double2 *A, int lda, const double2 *B, int ldb, double2 *C,
                       int ldc, hipStream_t s)
{
    const dim3 grid((cols + 63) / 64, (rows + 63) / 64);
    zgemm_seqk_kernel<<<grid, 256, 0, s>>>(rows, cols, K, A, lda, B, ldb, C, ldc);
}

__global__ __launch_bounds__(256) void robs_normalise_kernel(double2 *R, int d, int type)
{
    __shared__ double trace;
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < d; i++) t += R[(size_t)i * d + i].x;
        if (type == 2) t = t / double(d);
        trace = t;
    }
    __syncthreads();
    const double t = trace;
    for (size_t e = threadIdx.x; e < (size_t)d * d; e += blockDim.x) R[e] = make_double2(R[e].x / t, R[e].y / t);
}

void launch_robs_normalise(double2 *R, int d, int type, hipStream_t s)
{
    if (type > 0) robs_normalise_kernel<<<1, 256, 0, s>>>(R, d, type);
}

// ---------------------------------------------------------------------------
// Random observables of many trajectories: pn = phi / ||phi|| (the norm summed in
// ascending order on one lane), y = R pn as a GEMM over the trajectories, then
// <pn|y> in ascending order per trajectory (get_random_obs).
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline double2 state_at(const T *col, int dp, int k)
{
    return make_double2((double)col[k], (double)col[dp + k]);
}

template <typename T>
__global__ __launch_bounds__(64) void robs_pn_kernel(const T *slot0, const T *slot1, const TrajState *st, int tr0,
                                                     int n, int d, int dp, int L, double2 *pn)
{
    __shared__ double s_nrm;   // ||phi||
    const int c = blockIdx.x;
    const int tr = tr0 + c;
    const T *col = (st[tr].cur ? slot1 : slot0) + (size_t)tr * L;
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int k = 0; k < d; k++) {
            const double2 v = state_at(col, dp, k);
            s += v.x * v.x + v.y * v.y;
        }
        s_nrm = sqrt(s);
    }
    __syncthreads();
    const double nn = s_nrm;
    for (int k = threadIdx.x; k < d; k += 64) {
        const double2 v = state_at(col, dp, k);
        pn[(size_t)k * n + c] = make_double2(v.x / nn, v.y / nn);
    }
}

__global__ __launch_bounds__(256) void robs_dot_kernel(const double2 *pn, const double2 *y, int n, int d, int o,
                                                       int num, double2 *out, int tr0)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double rr = 0.0, ri = 0.0;
    for (int i = 0; i < d; i++) {
        const double2 p = pn[(size_t)i * n + c], m = y[(size_t)i * n + c];
        const double cr = p.x, ci = -p.y;
        rr = rr + (cr * m.x - ci * m.y);
        ri = ri + (ci * m.x + cr * m.y);
    }
    out[(size_t)(tr0 + c) * num + o] = make_double2(rr, ri);
}

void launch_random_obs(int prec, const void *slot0, const void *slot1, const TrajState *st, int tr0, int n, int d,
                       int dp, int L, const double2 *R, int num, double2 *pn, double2 *y, double2 *out,
                       hipStream_t s)
{
    if (n <= 0) return;
    if (prec == kFP64)
        robs_pn_kernel<double><<<n, 64, 0, s>>>((const double *)slot0, (const double *)slot1, st, tr0, n, d, dp, L,
                                                pn);
    else
        robs_pn_kernel<float><<<n, 64, 0, s>>>((const float *)slot0, (const float *)slot1, st, tr0, n, d, dp, L, pn);
    for (int o = 0; o < num; o++) {
        launch_zgemm_seqk(d, n, d, R + (size_t)o * d * d, d, pn, n, y, n, s);
        robs_dot_kernel<<<(n + 255) / 256, 256, 0, s>>>(pn, y, n, d, o, num, out, tr0);
    }
}

// ---------------------------------------------------------------------------
// LPN: one workgroup per trajectory or pair.  A state in work is its d entries
// as double2; lane 0 runs every sequential reduction.
// ---------------------------------------------------------------------------
constexpr int kLpnThreads = 256;

__device__ inline void sync() { __syncthreads(); }

// ||f||^2 in ascending order (norm_square, experiment.cpp:1263-1271): lane 0.
__device__ inline double seq_norm(const double2 *f, int d)
{
    double s = 0.0;
    for (int k = 0; k < d; k++) s += f[k].x * f[k].x + f[k].y * f[k].y;
    return s;
}

// Workgroup evaluation of one state f (d entries, in LDS or global memory):
// abs_diag_rho, mean and energy (get_mean_simple, get_energy on the band of H and
// H_drv: the dense loop's other terms are exact zeros), optionally norm and m2
// (with mean_start for the dispersion), and the random observables.  pn and mult
// are d-entry LDS scratch.  Results are valid on lane 0 (scalars) and in out arrays.
struct CharsOut {
    double norm, mean, energy, m2;
};

__device__ CharsOut eval_chars(const LpnArgs &a, const double2 *f, double *adr, double2 *obs, double2 *pn,
                               double2 *mult, bool with_m2)
{
    __shared__ double sh[2];
    const int d = a.d;
    CharsOut r{0.0, 0.0, 0.0, 0.0};
    if (threadIdx.x == 0) {
        const double nrm = seq_norm(f, d);
        double mn = 0.0;
        for (int k = 0; k < d; k++) mn += double(k) * ((f[k].x * f[k].x + f[k].y * f[k].y) * (1.0 / nrm));
        const double sn = sqrt(nrm);
        double en = 0.0;
        for (int i = 0; i < d; i++) {
            double sre = 0.0, sim = 0.0;
            const int j0 = i > 0 ? i - 1 : 0, j1 = i + 1 < d ? i + 1 : d - 1;
            for (int j = j0; j <= j1; j++) {
                const double h = i == j ? a.Hdiag[i] : a.Hoff[i < j ? i : j];
                const double hv = h + a.E0 * (i == j ? a.Hdrv[i] : 0.0);
                sre += hv * f[j].x / sn - 0.0 * f[j].y / sn;
                sim += hv * f[j].y / sn + 0.0 * f[j].x / sn;
            }
            en += f[i].x / sn * sre + f[i].y / sn * sim;
        }
        r.norm = nrm;
        r.mean = mn;
        r.energy = en;
        sh[0] = nrm;
        sh[1] = sqrt(nrm);
    }
    sync();
    const double nrm = sh[0], nn = sh[1];
    for (int k = threadIdx.x; k < d; k += blockDim.x) {
        adr[k] = (f[k].x * f[k].x + f[k].y * f[k].y) * (1.0 / nrm);
        pn[k] = make_double2(f[k].x / nn, f[k].y / nn);
    }
    sync();
    if (with_m2 && threadIdx.x == 0) {
        double mm2 = 0.0;
        for (int k = 0; k < d; k++) mm2 += (double(k) - r.mean) * (double(k) - r.mean) * adr[k];
        r.m2 = mm2;
    }
    for (int o = 0; o < a.num; o++) {
        const double2 *R = a.R + (size_t)o * d * d;
        for (int i = threadIdx.x; i < d; i += blockDim.x) {
            double re = 0.0, im = 0.0;
            for (int j = 0; j < d; j++) {
                const double2 q = R[(size_t)i * d + j];
                re = re + (q.x * pn[j].x - q.y * pn[j].y);
                im = im + (q.x * pn[j].y + q.y * pn[j].x);
            }
            mult[i] = make_double2(re, im);
        }
        sync();
        if (threadIdx.x == 0) {
            double rr = 0.0, ri = 0.0;
            for (int i = 0; i < d; i++) {
                const double cr = pn[i].x, ci = -pn[i].y;
                rr = rr + (cr * mult[i].x - ci * mult[i].y);
                ri = ri + (ci * mult[i].x + cr * mult[i].y);
            }
            obs[o] = make_double2(rr, ri);
        }
        sync();
    }
    return r;
}

template <typename T>
__device__ inline void load_state(const LpnArgs &a, int tr, double2 *f)
{
    const TrajState &s = a.st[tr];
    const T *col = (const T *)(s.cur ? a.slot1 : a.slot0) + (size_t)tr * a.L;
    for (int k = threadIdx.x; k < a.d; k += blockDim.x) f[k] = make_double2((double)col[k], (double)col[a.dp + k]);
}

template <typename T>
__device__ inline void store_state(const LpnArgs &a, int tr, const double2 *f)
{
    const TrajState &s = a.st[tr];
    T *col = (T *)(s.cur ? a.slot1 : a.slot0) + (size_t)tr * a.L;
    for (int k = threadIdx.x; k < a.d; k += blockDim.x) {
        col[k] = (T)f[k].x;
        col[a.dp + k] = (T)f[k].y;
    }
}

// LDS layout of the LPN kernels: three d-entry double2 arrays (state, pn, mult).
__device__ inline double2 *lds_arrays(int d, double2 **pn, double2 **mult)
{
    extern __shared__ double2 lds[];
    *pn = lds + d;
    *mult = lds + 2 * d;
    return lds;
}

// calc_chars_std(_start): norm, abs_diag_rho, mean, dispersion, m2, energy, random obs.
template <typename T>
__global__ __launch_bounds__(kLpnThreads) void lpn_chars_std_kernel(LpnArgs a, int tr0, int is_start)
{
    const int tr = tr0 + blockIdx.x;
    double2 *pn, *mult;
    double2 *f = lds_arrays(a.d, &pn, &mult);
    load_state<T>(a, tr, f);
    sync();
    const CharsOut c = eval_chars(a, f, a.adr + (size_t)tr * a.d, a.obs + (size_t)tr * a.num, pn, mult, true);
    if (threadIdx.x == 0) {
        a.norm[tr] = c.norm;
        a.mean[tr] = c.mean;
        a.energy[tr] = c.energy;
        if (is_start) a.mean_start[tr] = c.mean;
        const double ms = a.mean_start[tr];
        a.disp[tr] = (c.mean - ms) * (c.mean - ms);
        a.m2[tr] = c.m2;
    }
}

// calc_chars_lpn: abs_diag_rho, mean, energy and random observables into the lpn arrays.
template <typename T>
__global__ __launch_bounds__(kLpnThreads) void lpn_chars_lpn_kernel(LpnArgs a, int tr0)
{
    const int tr = tr0 + blockIdx.x;
    double2 *pn, *mult;
    double2 *f = lds_arrays(a.d, &pn, &mult);
    load_state<T>(a, tr, f);
    sync();
    const CharsOut c = eval_chars(a, f, a.adr + (size_t)tr * a.d, a.obs_lpn + (size_t)tr * a.num, pn, mult, false);
    if (threadIdx.x == 0) {
        a.mean_lpn[tr] = c.mean;
        a.energy_lpn[tr] = c.energy;
    }
}

// calc_delta_s (from the lpn values) / calc_delta_f (from the std values).
__device__ inline double lpn_delta(const LpnArgs &a, int tr, int base, bool from_lpn)
{
    const double *mn = from_lpn ? a.mean_lpn : a.mean;
    const double *en = from_lpn ? a.energy_lpn : a.energy;
    const double2 *ob = from_lpn ? a.obs_lpn : a.obs;
    if (a.lpn_type == -1) return fabs(mn[tr] - mn[base]) / double(a.d);
    if (a.lpn_type == -2) return fabs(en[tr] - en[base]) / a.max_energy;
    const double2 b = ob[(size_t)base * a.num + a.lpn_type], v = ob[(size_t)tr * a.num + a.lpn_type];
    const double dr = b.x - v.x, di = b.y - v.y;
    return sqrt(dr * dr + di * di);
}

// var_trajectory_lpn: the partner's state = the original's + eps w, renormalised
// to the original's norm, eps bisected in log10 until the lpn distance is
// lpn_delta_s within lpn_eps_error (lpn_eps_deep halvings at most).  Leaves the
// partner's lpn values set (the last trial's calc_chars_lpn) and its state in f
// (double) and in its slot.  The bisection's bounds live on lane 0; the other
// lanes read eps and the stop flag from LDS.
template <typename T>
__device__ void perturb(const LpnArgs &a, int tr, int base, double2 *f, double2 *pn, double2 *mult)
{
    __shared__ double sh[3];
    __shared__ double s_eps;
    __shared__ int s_stop;
    const int d = a.d;
    double2 *w = a.work + (size_t)(tr - a.half) * 2 * d;   // direction
    double2 *orig = w + d;                                  // the original's state
    load_state<T>(a, base, orig);
    if (threadIdx.x == 0) {
        uint64_t x = a.var_x[tr];
        // 2d uniforms on [-1, 1): the real parts, then the imaginary parts
        for (int i = 0; i < d; i++) w[i].x = mcg31_next(x, a.var_stride, -1.0, 1.0);
        for (int i = 0; i < d; i++) w[i].y = mcg31_next(x, a.var_stride, -1.0, 1.0);
        a.var_x[tr] = x;
    }
    __threadfence_block();
    sync();
    if (threadIdx.x == 0) {
        sh[0] = seq_norm(w, d);
        sh[1] = seq_norm(orig, d);
    }
    sync();
    const double nv2 = sh[0], n2 = sh[1];
    for (int i = threadIdx.x; i < d; i += blockDim.x) w[i] = make_double2(w[i].x / sqrt(nv2), w[i].y / sqrt(nv2));
    double lo = a.eps_low, hi = a.eps_high, diff = hi - lo;   // lane 0's copies are the ones used
    int deep = 0;
    bool ran = false;
    if (threadIdx.x == 0) {
        s_eps = lpn_exp10((lo + hi) * 0.5);
        s_stop = !(deep < a.eps_deep);
    }
    __threadfence_block();
    sync();
    while (!s_stop) {
        ran = true;
        const double eps = s_eps;
        for (int i = threadIdx.x; i < d; i += blockDim.x)
            f[i] = make_double2(orig[i].x + eps * w[i].x, orig[i].y + eps * w[i].y);
        sync();
        if (threadIdx.x == 0) sh[2] = seq_norm(f, d);
        sync();
        const double nm2 = sh[2];
        for (int i = threadIdx.x; i < d; i += blockDim.x) {
            const double2 v = make_double2(f[i].x / sqrt(nm2), f[i].y / sqrt(nm2));
            f[i] = make_double2(v.x * sqrt(n2), v.y * sqrt(n2));
        }
        sync();
        const CharsOut c =
            eval_chars(a, f, a.adr + (size_t)tr * d, a.obs_lpn + (size_t)tr * a.num, pn, mult, false);
        if (threadIdx.x == 0) {
            a.mean_lpn[tr] = c.mean;
            a.energy_lpn[tr] = c.energy;
            const double ds = lpn_delta(a, tr, base, true);
            double e = eps;
            int stop = 0;
            if (fabs(ds - a.delta_s_target) < a.eps_error) {
                stop = 1;
            } else {
                if (ds > a.delta_s_target) {
                    hi -= 0.5 * diff;
                    diff = hi - lo;
                    deep++;
                    e = lpn_exp10((lo + hi) * 0.5);
                }
                if (ds < a.delta_s_target) {
                    lo += 0.5 * diff;
                    diff = hi - lo;
                    deep++;
                    e = lpn_exp10((lo + hi) * 0.5);
                }
                if (!(ds > a.delta_s_target) && !(ds < a.delta_s_target)) stop = 1;   // equal or NaN
            }
            s_eps = e;
            s_stop = stop || !(deep < a.eps_deep);
        }
        __threadfence_block();
        sync();
    }
    if (ran) store_state<T>(a, tr, f);
}

// calc_chars_lpn_start: the lpn values start as the std values, lambda at 0, and
// delta_s is the std distance to the original.
__global__ void lpn_start_kernel(LpnArgs a, int tr0, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int tr = tr0 + i;
    const int base = tr >= a.half ? tr - a.half : tr;
    a.lambda[tr] = 0.0;
    a.lambda_now[tr] = 0.0;
    a.mean_lpn[tr] = a.mean[tr];
    a.energy_lpn[tr] = a.energy[tr];
    a.delta_s[tr] = lpn_delta(a, tr, base, false);
    for (int o = 0; o < a.num; o++) a.obs_lpn[(size_t)tr * a.num + o] = a.obs[(size_t)tr * a.num + o];
}

// copy_trajectory_lpn + copy_stream_lpn: state, RNG stream and eta of the original.
template <typename T>
__global__ __launch_bounds__(kLpnThreads) void lpn_pair_start_kernel(LpnArgs a)
{
    const int tr = a.half + blockIdx.x, base = blockIdx.x;
    double2 *pn, *mult;
    double2 *f = lds_arrays(a.d, &pn, &mult);
    if (threadIdx.x == 0) {
        a.st[tr].rng = a.st[base].rng;
        a.st[tr].eta = a.st[base].eta;
    }
    load_state<T>(a, base, f);
    sync();
    store_state<T>(a, tr, f);
    sync();
    perturb<T>(a, tr, base, f, pn, mult);
}

// One period's pair step (experiment.cpp:218-252 with lambda_lpn / lambda_lpn_now).
template <typename T>
__global__ __launch_bounds__(kLpnThreads) void lpn_pairs_kernel(LpnArgs a, double curr_time)
{
    __shared__ int renorm;
    __shared__ double df_sh;
    const int tr = a.half + blockIdx.x, base = blockIdx.x;
    double2 *pn, *mult;
    double2 *f = lds_arrays(a.d, &pn, &mult);
    if (threadIdx.x == 0) {
        const double df = lpn_delta(a, tr, base, false);
        const double thr = a.df_high > a.delta_s[tr] * 10.0 ? a.df_high : a.delta_s[tr] * 10.0;
        const int rn = a.lambda_per_periods != 0 || df > thr || df < a.df_low;
        if (a.save) a.rec_flag[tr] = rn;
        if (rn) {
            a.lambda[tr] += lpn_log(df / a.delta_s[tr] + 1.0e-16);
            a.num_renorms[tr] += 1;
            if (a.save) {
                a.rec_ratio[tr] = df / a.delta_s[tr];
                a.rec_ds[tr] = a.delta_s[tr];
                a.rec_df[tr] = df;
            }
            a.st[tr].rng = a.st[base].rng;
            a.st[tr].eta = a.st[base].eta;
        }
        renorm = rn;
        df_sh = df;
    }
    sync();
    if (renorm) {
        perturb<T>(a, tr, base, f, pn, mult);
        sync();
    } else {
        load_state<T>(a, tr, f);
        sync();
    }
    // calc_chars_lpn of the (renormalised) partner
    const CharsOut c = eval_chars(a, f, a.adr + (size_t)tr * a.d, a.obs_lpn + (size_t)tr * a.num, pn, mult, false);
    if (threadIdx.x == 0) {
        a.mean_lpn[tr] = c.mean;
        a.energy_lpn[tr] = c.energy;
    }
    __threadfence_block();
    sync();
    if (threadIdx.x == 0) {
        if (renorm) {
            a.delta_s[tr] = lpn_delta(a, tr, base, true);
            a.lambda_now[tr] = a.lambda[tr] / curr_time;
        } else {
            a.lambda_now[tr] = (a.lambda[tr] + lpn_log(df_sh / a.delta_s[tr] + 1.0e-16)) / curr_time;
        }
    }
}

static size_t lpn_lds(const LpnArgs &a) { return (size_t)3 * a.d * sizeof(double2); }

template <typename K>
static void allow_lds(K kernel, size_t bytes)
{
    if (bytes > 64 * 1024) (void)hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

void launch_lpn_chars_std(const LpnArgs &a, int tr0, int n, int is_start, hipStream_t s)
{
    if (n <= 0) return;
    const size_t lds = lpn_lds(a);
    if (a.prec == kFP64) {
        allow_lds(lpn_chars_std_kernel<double>, lds);
        lpn_chars_std_kernel<double><<<n, kLpnThreads, lds, s>>>(a, tr0, is_start);
    } else {
        allow_lds(lpn_chars_std_kernel<float>, lds);
        lpn_chars_std_kernel<float><<<n, kLpnThreads, lds, s>>>(a, tr0, is_start);
    }
}

void launch_lpn_chars_lpn(const LpnArgs &a, int tr0, int n, hipStream_t s)
{
    if (n <= 0) return;
    const size_t lds = lpn_lds(a);
    if (a.prec == kFP64) {
        allow_lds(lpn_chars_lpn_kernel<double>, lds);
        lpn_chars_lpn_kernel<double><<<n, kLpnThreads, lds, s>>>(a, tr0);
    } else {
        allow_lds(lpn_chars_lpn_kernel<float>, lds);
        lpn_chars_lpn_kernel<float><<<n, kLpnThreads, lds, s>>>(a, tr0);
    }
}

void launch_lpn_start(const LpnArgs &a, int tr0, int n, hipStream_t s)
{
    if (n > 0) lpn_start_kernel<<<(n + 255) / 256, 256, 0, s>>>(a, tr0, n);
}

void launch_lpn_pair_start(const LpnArgs &a, hipStream_t s)
{
    const size_t lds = lpn_lds(a);
    if (a.prec == kFP64) {
        allow_lds(lpn_pair_start_kernel<double>, lds);
        lpn_pair_start_kernel<double><<<a.half, kLpnThreads, lds, s>>>(a);
    } else {
        allow_lds(lpn_pair_start_kernel<float>, lds);
        lpn_pair_start_kernel<float><<<a.half, kLpnThreads, lds, s>>>(a);
    }
}

void launch_lpn_pairs(const LpnArgs &a, double curr_time, hipStream_t s)
{
    const size_t lds = lpn_lds(a);
    if (a.prec == kFP64) {
        allow_lds(lpn_pairs_kernel<double>, lds);
        lpn_pairs_kernel<double><<<a.half, kLpnThreads, lds, s>>>(a, curr_time);
    } else {
        allow_lds(lpn_pairs_kernel<float>, lds);
        lpn_pairs_kernel<float><<<a.half, kLpnThreads, lds, s>>>(a, curr_time);
    }
}

}  // namespace qjx

// ---------------------------------------------------------------------------
// Host side: the random observable matrices and the STD path's random observables.
// ---------------------------------------------------------------------------
namespace {

uint64_t mcg31_pow(uint64_t a, uint64_t e)
{
    uint64_t r = 1;
    a %= qjx::kMcg31M;
    while (e) {
        if (e & 1ULL) r = r * a % qjx::kMcg31M;
        a = a * a % qjx::kMcg31M;
        e >>= 1;
    }
    return r;
}

const double kPiObs = 3.1415926535897932384626433832795;   // reference config.h PI

}  // namespace

namespace qjx {

void mcg31_stream(uint64_t seed, int k, int n, uint64_t *x, uint64_t *stride)
{
    uint64_t x0 = seed % kMcg31M;
    if (x0 == 0) x0 = 1;
    *x = x0 * mcg31_pow(1132489760ULL, (uint64_t)k) % kMcg31M;
    *stride = mcg31_pow(1132489760ULL, (uint64_t)n);
}

// init_random_obs (newdel.cpp:3652-3722): X = (G_re + i G_im) / 2 from d^2 + d^2
// Box-Muller Gaussians of MCG31(3873247) leap-frog(seed + o, mns) (drawn here, with
// the host C library's log and sin), R = X X^dag on the device in ascending k,
// divided by Tr R (type 1) or Tr R / d (type 2).
int build_random_obs(int d, int num, int seed, int mns, int type, double2 **out, hipStream_t s)
{
    *out = nullptr;
    if (num <= 0) return QJX_OK;
    const size_t dd = (size_t)d * d;
    double2 *R = nullptr, *X = nullptr, *Xh = nullptr;
    if (hipMalloc((void **)&R, num * dd * sizeof(double2)) != hipSuccess ||
        hipMalloc((void **)&X, dd * sizeof(double2)) != hipSuccess ||
        hipMalloc((void **)&Xh, dd * sizeof(double2)) != hipSuccess) {
        if (R) (void)hipFree(R);
        if (X) (void)hipFree(X);
        if (Xh) (void)hipFree(Xh);
        set_last_error("random observable matrices do not fit in device memory");
        return QJX_ERR_NOMEM;
    }
    std::vector<double2> hx(dd), hxh(dd);
    std::vector<double> gr(dd), gi(dd);
    int rc = QJX_OK;
    for (int o = 0; o < num && rc == QJX_OK; o++) {
        uint64_t x, stride;
        mcg31_stream(3873247ULL, seed + o, mns, &x, &stride);
        auto gauss = [&]() {
            const double u1 = mcg31_next(x, stride, 0.0, 1.0);
            const double u2 = mcg31_next(x, stride, 0.0, 1.0);
            return std::sqrt(-2.0 * std::log(u1)) * std::sin(2.0 * kPiObs * u2);
        };
        for (size_t i = 0; i < dd; i++) gr[i] = gauss();
        for (size_t i = 0; i < dd; i++) gi[i] = gauss();
        for (int i = 0; i < d; i++)
            for (int k = 0; k < d; k++) {
                const size_t e = (size_t)i * d + k;
                hx[e] = make_double2(0.5 * gr[e], 0.5 * gi[e]);
                hxh[(size_t)k * d + i] = make_double2(0.5 * gr[e], -(0.5 * gi[e]));   // X^dag
            }
        if (hipMemcpyAsync(X, hx.data(), dd * sizeof(double2), hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(Xh, hxh.data(), dd * sizeof(double2), hipMemcpyHostToDevice, s) != hipSuccess) {
            rc = QJX_ERR_HIP;
            break;
        }
        launch_zgemm_seqk(d, d, d, X, d, Xh, d, R + o * dd, d, s);
        launch_robs_normalise(R + o * dd, d, type, s);
        if (hipStreamSynchronize(s) != hipSuccess) rc = QJX_ERR_HIP;   // hx / hxh are reused
    }
    (void)hipFree(X);
    (void)hipFree(Xh);
    if (rc != QJX_OK) {
        (void)hipFree(R);
        set_last_error("random observable build failed");
        return rc;
    }
    *out = R;
    return QJX_OK;
}

}  // namespace qjx

// calc_random_obs on the STD path (core.cpp:694, 724, 12614-12622): the context keeps
// the matrices of the last (num, seed, mns, type) asked for.
extern "C" int qjx_random_obs(qjx_ctx *c, int num, int seed, int mns, int type, double *out)
{
    if (!c || !c->have_ens) {
        qjx::set_last_error("ensemble not initialised");
        return QJX_ERR_ARG;
    }
    if (num < 1 || mns < 1 || type < 0 || type > 2) {
        qjx::set_last_error("qjx_random_obs: num >= 1, mns >= 1 and type 0, 1 or 2 are required");
        return QJX_ERR_ARG;
    }
    if (!c->robs_R || c->robs_num != num || c->robs_seed != seed || c->robs_mns != mns || c->robs_type != type) {
        if (c->robs_R) (void)hipFree(c->robs_R);
        c->robs_R = nullptr;
        double2 *R = nullptr;
        const int rc = qjx::build_random_obs(c->d, num, seed, mns, type, &R, c->stream);
        if (rc) return rc;
        c->robs_R = R;
        c->robs_num = num;
        c->robs_seed = seed;
        c->robs_mns = mns;
        c->robs_type = type;
    }
    // trajectories in chunks whose scratch (pn, y: d x chunk each) stays near 256 MiB
    const int d = c->d, M = c->M;
    const int chunk = std::max(1, std::min(M, (int)((size_t(256) << 20) / ((size_t)2 * d * sizeof(double2)))));
    const size_t need = (size_t)2 * d * chunk + (size_t)M * num;
    if (c->robs_scratch_n < need) {
        if (c->robs_scratch) (void)hipFree(c->robs_scratch);
        c->robs_scratch = nullptr;
        c->robs_scratch_n = 0;
        if (hipMalloc((void **)&c->robs_scratch, need * sizeof(double2)) != hipSuccess) {
            qjx::set_last_error("random observable scratch does not fit in device memory");
            return QJX_ERR_NOMEM;
        }
        c->robs_scratch_n = need;
    }
    double2 *pn = c->robs_scratch, *y = pn + (size_t)d * chunk, *res = y + (size_t)d * chunk;
    for (int tr0 = 0; tr0 < M; tr0 += chunk)
        qjx::launch_random_obs(c->prec, c->slot[0], c->slot[1], c->st, tr0, std::min(chunk, M - tr0), d, c->dp, c->L,
                               c->robs_R, num, pn, y, res, c->stream);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(out, res, (size_t)M * num * sizeof(double2), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        qjx::set_last_error("random observable launch failed");
        return QJX_ERR_HIP;
    }
    return QJX_OK;
}
