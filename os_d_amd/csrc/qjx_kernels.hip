// qjx_kernels.hip -- HIP kernels of the MI355X trajectory engine (gfx950) outside
// the propagation iteration (qjx_step.hip): ensemble initialisation, observables,
// state export/import and the propagator build (dense complex FP64 GEMM,
// Taylor-polynomial helpers, block-real packing).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "qjx_device.h"
#include "qjx_internal.h"
#include "qjx_kernels.h"

namespace qjx {

// ---------------------------------------------------------------------------
// Ensemble initialisation (newdel_exp.cpp:332-375, 656-679, 244-296)
// ---------------------------------------------------------------------------
__global__ void init_states_kernel(TrajState *st, int M, int seed)
{
    const int tr = blockIdx.x * blockDim.x + threadIdx.x;
    if (tr >= M) return;
    TrajState s;
    s.t = 0.0;
    s.eta = 0.0;
    s.rng = mcg_mul(777ULL, mcg_pow(302875106592253ULL, (uint64_t)(seed + tr)));
    s.cnt0 = 0;
    s.bits = 0;
    s.lvl = 0;
    s.b = 0;
    s.pos = 0;
    s.target = 0;
    s.cur = 0;
    s.njumps = 0;
    s.logn = 0;
    s.pad = 0;
    st[tr] = s;
}

template <typename T>
__global__ void init_psi_kernel(T *slot0, T *slot1, int M, int d, int dp, int L, int start_type,
                                int start_state)
{
    const size_t total = (size_t)M * L;
    const T uni = (T)sqrt(1.0 / double(d));
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
         i += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(i % L);
        T v = T(0);
        if (start_type == 0) {
            if (k == start_state) v = T(1);
        } else if (k < d) {
            v = uni;
        }
        slot0[i] = v;
        slot1[i] = T(0);
    }
}

void launch_init_ensemble(int prec, void *slot0, void *slot1, TrajState *st, int M, int d, int dp,
                          int L, int seed, int start_type, int start_state, hipStream_t s)
{
    init_states_kernel<<<(M + 255) / 256, 256, 0, s>>>(st, M, seed);
    const int blocks = 2048;
    if (prec == kFP64)
        init_psi_kernel<double><<<blocks, 256, 0, s>>>((double *)slot0, (double *)slot1, M, d, dp, L,
                                                        start_type, start_state);
    else
        init_psi_kernel<float><<<blocks, 256, 0, s>>>((float *)slot0, (float *)slot1, M, d, dp, L,
                                                       start_type, start_state);
}

__global__ void draw_eta_kernel(TrajState *st, int M, uint64_t stride)
{
    const int tr = blockIdx.x * blockDim.x + threadIdx.x;
    if (tr >= M) return;
    TrajState s = st[tr];
    double e = 0.0;
    while (e == 0.0) e = mcg_next(s.rng, stride);
    s.eta = e;
    st[tr] = s;
}

void launch_draw_eta(TrajState *st, int M, uint64_t stride, hipStream_t s)
{
    draw_eta_kernel<<<(M + 255) / 256, 256, 0, s>>>(st, M, stride);
}

__global__ void reset_times_kernel(TrajState *st, int M)
{
    const int tr = blockIdx.x * blockDim.x + threadIdx.x;
    if (tr < M) st[tr].t = 0.0;
}

void launch_reset_times(TrajState *st, int M, hipStream_t s)
{
    reset_times_kernel<<<(M + 255) / 256, 256, 0, s>>>(st, M);
}

__global__ void reset_logn_kernel(TrajState *st, int M)
{
    const int tr = blockIdx.x * blockDim.x + threadIdx.x;
    if (tr < M) st[tr].logn = 0;
}

void launch_reset_logn(TrajState *st, int M, hipStream_t s)
{
    reset_logn_kernel<<<(M + 255) / 256, 256, 0, s>>>(st, M);
}

// ---------------------------------------------------------------------------
// Observables (core.cpp:665-725, experiment.cpp:1395-1469), one wave per
// trajectory, FP64 arithmetic.  H + E*H_drv is real tridiagonal: Hdiag holds
// the diagonal with E*H_drv folded in, Hoff[k] = H_{k,k+1} = H_{k+1,k}.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void chars_kernel(const T *slot0, const T *slot1,
                                                    const TrajState *st, int M, int d, int dp,
                                                    int L, const double *Hdiag,
                                                    const double *Hoff, int is_start,
                                                    double *mean_start, double *norm,
                                                    double *mean, double *disp, double *m2,
                                                    double *energy, double *adr)
{
    const int tr = blockIdx.x * 4 + threadIdx.x / 64;
    const int lane = threadIdx.x & 63;
    if (tr >= M) return;
    const T *x = (st[tr].cur ? slot1 : slot0) + (size_t)tr * L;
    double n = 0.0;
    for (int k = lane; k < d; k += 64) {
        const double re = (double)x[k], im = (double)x[dp + k];
        n += re * re + im * im;
    }
    n = wave_sum(n);
    const double inv = 1.0 / n;
    double mn = 0.0;
    for (int k = lane; k < d; k += 64) {
        const double re = (double)x[k], im = (double)x[dp + k];
        const double a = (re * re + im * im) * inv;
        mn += double(k) * a;
        if (adr) adr[(size_t)tr * d + k] = a;
    }
    mn = wave_sum(mn);
    double mm2 = 0.0, en = 0.0;
    const double sn = sqrt(n);
    for (int k = lane; k < d; k += 64) {
        const double re = (double)x[k], im = (double)x[dp + k];
        const double a = (re * re + im * im) * inv;
        mm2 += (double(k) - mn) * (double(k) - mn) * a;
        // sm_k = sum_j H_kj psi_j / sqrt(n), j ascending
        double sre = 0.0, sim = 0.0;
        if (k > 0) {
            sre += Hoff[k - 1] * ((double)x[k - 1] / sn);
            sim += Hoff[k - 1] * ((double)x[dp + k - 1] / sn);
        }
        sre += Hdiag[k] * (re / sn);
        sim += Hdiag[k] * (im / sn);
        if (k + 1 < d) {
            sre += Hoff[k] * ((double)x[k + 1] / sn);
            sim += Hoff[k] * ((double)x[dp + k + 1] / sn);
        }
        en += re / sn * sre + im / sn * sim;
    }
    mm2 = wave_sum(mm2);
    en = wave_sum(en);
    if (lane == 0) {
        if (is_start) mean_start[tr] = mn;
        const double ms = mean_start[tr];
        norm[tr] = n;
        mean[tr] = mn;
        disp[tr] = (mn - ms) * (mn - ms);
        m2[tr] = mm2;
        energy[tr] = en;
    }
}

void launch_chars(int prec, const void *slot0, const void *slot1, const TrajState *st, int M,
                  int d, int dp, int L, const double *Hdiag, const double *Hoff, int is_start,
                  double *mean_start, double *norm, double *mean, double *disp, double *m2,
                  double *energy, double *adr, hipStream_t s)
{
    const int blocks = (M + 3) / 4;
    if (prec == kFP64)
        chars_kernel<double><<<blocks, 256, 0, s>>>((const double *)slot0, (const double *)slot1, st,
                                                     M, d, dp, L, Hdiag, Hoff, is_start, mean_start,
                                                     norm, mean, disp, m2, energy, adr);
    else
        chars_kernel<float><<<blocks, 256, 0, s>>>((const float *)slot0, (const float *)slot1, st, M,
                                                    d, dp, L, Hdiag, Hoff, is_start, mean_start,
                                                    norm, mean, disp, m2, energy, adr);
}

// adr_avg: partial sums over fixed chunks of trajectories, then the chunks in
// order; each term is adr / M as in dump.cpp:50-56.
constexpr int kAdrChunks = 64;

__global__ void adr_partial_kernel(const double *adr, int M, int d, double *part)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (k >= d) return;
    const int per = (M + kAdrChunks - 1) / kAdrChunks;
    const int t0 = c * per, t1 = min(M, t0 + per);
    double s = 0.0;
    for (int tr = t0; tr < t1; tr++) s += adr[(size_t)tr * d + k] / double(M);
    part[(size_t)c * d + k] = s;
}

__global__ void adr_final_kernel(const double *part, int d, double *out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d) return;
    double s = 0.0;
    for (int c = 0; c < kAdrChunks; c++) s += part[(size_t)c * d + k];
    out[k] = s;
}

void launch_adr_avg(const double *adr, int M, int d, double *out, hipStream_t s)
{
    // `out` must hold d + kAdrChunks * d doubles; the tail is scratch.
    double *part = out + d;
    dim3 grid((d + 63) / 64, kAdrChunks);
    adr_partial_kernel<<<grid, 64, 0, s>>>(adr, M, d, part);
    adr_final_kernel<<<(d + 63) / 64, 64, 0, s>>>(part, d, out);
}

template <typename T>
__global__ void export_state_kernel(const T *slot0, const T *slot1, const TrajState *st, int M,
                                    int d, int dp, int L, double *out)
{
    const size_t total = (size_t)M * d;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
         i += (size_t)gridDim.x * blockDim.x) {
        const int tr = (int)(i / d), k = (int)(i % d);
        const T *x = (st[tr].cur ? slot1 : slot0) + (size_t)tr * L;
        out[2 * i] = (double)x[k];
        out[2 * i + 1] = (double)x[dp + k];
    }
}

template <typename T>
__global__ void import_state_kernel(T *slot0, T *slot1, const TrajState *st, int M, int d, int dp,
                                    int L, const double *in)
{
    const size_t total = (size_t)M * d;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
         i += (size_t)gridDim.x * blockDim.x) {
        const int tr = (int)(i / d), k = (int)(i % d);
        T *x = (st[tr].cur ? slot1 : slot0) + (size_t)tr * L;
        x[k] = (T)in[2 * i];
        x[dp + k] = (T)in[2 * i + 1];
    }
}

void launch_export_state(int prec, const void *slot0, const void *slot1, const TrajState *st,
                         int M, int d, int dp, int L, double *out, hipStream_t s)
{
    if (prec == kFP64)
        export_state_kernel<double><<<1024, 256, 0, s>>>((const double *)slot0, (const double *)slot1,
                                                          st, M, d, dp, L, out);
    else
        export_state_kernel<float><<<1024, 256, 0, s>>>((const float *)slot0, (const float *)slot1,
                                                         st, M, d, dp, L, out);
}

void launch_import_state(int prec, void *slot0, void *slot1, const TrajState *st, int M, int d,
                         int dp, int L, const double *in, hipStream_t s)
{
    if (prec == kFP64)
        import_state_kernel<double><<<1024, 256, 0, s>>>((double *)slot0, (double *)slot1, st, M, d,
                                                          dp, L, in);
    else
        import_state_kernel<float><<<1024, 256, 0, s>>>((float *)slot0, (float *)slot1, st, M, d, dp,
                                                         L, in);
}

// ---------------------------------------------------------------------------
// Dense complex FP64 GEMM for the propagator build: C = A B, row-major n x n.
// 32x32 output tile per block, 2x2 outputs per thread, K tile of 16.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void zgemm_kernel(int n, const double2 *A, const double2 *B,
                                                    double2 *C)
{
    __shared__ double2 As[32][17];
    __shared__ double2 Bs[16][33];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int row0 = blockIdx.y * 32, col0 = blockIdx.x * 32;
    double2 acc[2][2];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) acc[i][j] = make_double2(0.0, 0.0);
    for (int k0 = 0; k0 < n; k0 += 16) {
        for (int e = threadIdx.x; e < 32 * 16; e += 256) {
            const int r = e / 16, c = e % 16;
            const int gr = row0 + r, gc = k0 + c;
            As[r][c] = (gr < n && gc < n) ? A[(size_t)gr * n + gc] : make_double2(0.0, 0.0);
            const int br = e / 32, bc = e % 32;
            const int gbr = k0 + br, gbc = col0 + bc;
            Bs[br][bc] = (gbr < n && gbc < n) ? B[(size_t)gbr * n + gbc] : make_double2(0.0, 0.0);
        }
        __syncthreads();
        for (int kk = 0; kk < 16; kk++) {
            double2 a0 = As[ty * 2][kk], a1 = As[ty * 2 + 1][kk];
            double2 b0 = Bs[kk][tx * 2], b1 = Bs[kk][tx * 2 + 1];
            acc[0][0].x += a0.x * b0.x - a0.y * b0.y;
            acc[0][0].y += a0.x * b0.y + a0.y * b0.x;
            acc[0][1].x += a0.x * b1.x - a0.y * b1.y;
            acc[0][1].y += a0.x * b1.y + a0.y * b1.x;
            acc[1][0].x += a1.x * b0.x - a1.y * b0.y;
            acc[1][0].y += a1.x * b0.y + a1.y * b0.x;
            acc[1][1].x += a1.x * b1.x - a1.y * b1.y;
            acc[1][1].y += a1.x * b1.y + a1.y * b1.x;
        }
        __syncthreads();
    }
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
            const int r = row0 + ty * 2 + i, c = col0 + tx * 2 + j;
            if (r < n && c < n) C[(size_t)r * n + c] = acc[i][j];
        }
}

void launch_zgemm(int n, const double *A, const double *B, double *C, hipStream_t s)
{
    dim3 grid((n + 31) / 32, (n + 31) / 32);
    zgemm_kernel<<<grid, 256, 0, s>>>(n, (const double2 *)A, (const double2 *)B, (double2 *)C);
}

__global__ void zaxpby_eye_kernel(int n, const double *A, double alpha, double beta, double *C)
{
    const size_t total = (size_t)n * n;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
         i += (size_t)gridDim.x * blockDim.x) {
        const size_t r = i / n, c = i % n;
        double re = A ? alpha * A[2 * i] : 0.0;
        const double im = A ? alpha * A[2 * i + 1] : 0.0;
        if (r == c) re += beta;
        C[2 * i] = re;
        C[2 * i + 1] = im;
    }
}

void launch_zaxpby_eye(int n, const double *A, double alpha, double beta, double *C, hipStream_t s)
{
    zaxpby_eye_kernel<<<1024, 256, 0, s>>>(n, A, alpha, beta, C);
}

__global__ void zaxpy_kernel(int n, const double *A, double alpha, double *C)
{
    const size_t total = (size_t)n * n * 2;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
         i += (size_t)gridDim.x * blockDim.x)
        C[i] += alpha * A[i];
}

void launch_zaxpy(int n, const double *A, double alpha, double *C, hipStream_t s)
{
    zaxpy_kernel<<<1024, 256, 0, s>>>(n, A, alpha, C);
}

// Block-real row-major operand: element (i, k) of [[Re U, -Im U], [Im U, Re U]]
// at i * 2dp + k; rows/columns beyond the n x n blocks are zero.
template <typename T>
__global__ void pack_propagator_kernel(int n, int dp, const double *U, T *dst)
{
    const int L = 2 * dp;
    const size_t total = (size_t)L * L;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(e / L), k = (int)(e % L);
        const int ib = i / dp, ii = i % dp, kb = k / dp, kk = k % dp;
        double v = 0.0;
        if (ii < n && kk < n) {
            const double re = U[2 * ((size_t)ii * n + kk)], im = U[2 * ((size_t)ii * n + kk) + 1];
            if (ib == kb) v = re;
            else if (ib == 0) v = -im;   // top-right block
            else v = im;                 // bottom-left block
        }
        dst[e] = (T)v;
    }
}

// 3 x BF16 form: hi = bf16(u), lo = bf16(u - hi) of the FP32-rounded element.
__global__ void pack_propagator_bf16x3_kernel(int n, int dp, const double *U, __bf16 *hi, __bf16 *lo)
{
    const int L = 2 * dp;
    const size_t total = (size_t)L * L;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(e / L), k = (int)(e % L);
        const int ib = i / dp, ii = i % dp, kb = k / dp, kk = k % dp;
        double v = 0.0;
        if (ii < n && kk < n) {
            const double re = U[2 * ((size_t)ii * n + kk)], im = U[2 * ((size_t)ii * n + kk) + 1];
            if (ib == kb) v = re;
            else if (ib == 0) v = -im;
            else v = im;
        }
        const float x = (float)v;
        const __bf16 h = (__bf16)x;
        hi[e] = h;
        lo[e] = (__bf16)(x - (float)h);
    }
}

// Complex-pair rows: element (i, k) of the dp x 2dp operand [Re U | Im U] at
// i * 2dp + k, zero beyond the n x n block; the 3 x BF16 form splits each FP32-rounded
// element into hi = bf16(u) and lo = bf16(u - hi).
template <bool SPLIT>
__global__ void pack_propagator_cp_kernel(int n, int dp, const double *U, void *dst, void *dst_lo)
{
    const int L = 2 * dp;
    const size_t total = (size_t)dp * L;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(e / L), k = (int)(e % L);
        const int kk = k < dp ? k : k - dp;
        double v = 0.0;
        if (i < n && kk < n) v = U[2 * ((size_t)i * n + kk) + (k < dp ? 0 : 1)];
        const float x = (float)v;
        if (SPLIT) {
            const __bf16 h = (__bf16)x;
            static_cast<__bf16 *>(dst)[e] = h;
            static_cast<__bf16 *>(dst_lo)[e] = (__bf16)(x - (float)h);
        } else {
            static_cast<float *>(dst)[e] = x;
        }
    }
}

bool cp_layout(int prec, int gemm_kind) { return prec != kFP64 && gemm_kind == kGemmMfma; }

void launch_pack_propagator(int prec, bool cp, int n, int dp, const double *U, void *dst, void *dst_lo,
                            hipStream_t s)
{
    if (cp && prec == kBF16x3) pack_propagator_cp_kernel<true><<<2048, 256, 0, s>>>(n, dp, U, dst, dst_lo);
    else if (cp) pack_propagator_cp_kernel<false><<<2048, 256, 0, s>>>(n, dp, U, dst, nullptr);
    else if (prec == kFP64) pack_propagator_kernel<double><<<2048, 256, 0, s>>>(n, dp, U, (double *)dst);
    else if (prec == kBF16x3)
        pack_propagator_bf16x3_kernel<<<2048, 256, 0, s>>>(n, dp, U, (__bf16 *)dst, (__bf16 *)dst_lo);
    else pack_propagator_kernel<float><<<2048, 256, 0, s>>>(n, dp, U, (float *)dst);
}

// ---------------------------------------------------------------------------
// Ensemble density matrix rho = (1/M) sum_tr psi psi^+ / ||psi||^2 and its purity
// Tr rho^2.  G = X W X^T over the block-real state columns X (L x M) with
// W = diag(1 / (M ||psi_tr||^2)) is accumulated as split-K partial sums of 64 x 64
// tiles (FP64, fixed order); rho is read off the four dp x dp blocks of G:
// Re rho = G_rr + G_ii, Im rho = G_ir - G_ri.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void density_weight_kernel(const T *slot0, const T *slot1,
                                                             const TrajState *st, int M, int L, double *w)
{
    const int tr = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (tr >= M) return;
    const T *x = (st[tr].cur ? slot1 : slot0) + (size_t)tr * L;
    double s = 0.0;
    for (int k = lane; k < L; k += 64) {
        const double v = (double)x[k];
        s += v * v;
    }
    s = wave_sum(s);
    if (lane == 0) w[tr] = s > 0.0 ? 1.0 / ((double)M * s) : 0.0;
}

constexpr int kDenTile = 64, kDenK = 16;

template <typename T>
__global__ __launch_bounds__(256) void density_gram_kernel(const T *slot0, const T *slot1,
                                                           const TrajState *st, const double *w, int M,
                                                           int L, int kper, double *part)
{
    __shared__ double As[kDenK][kDenTile + 1];
    __shared__ double Bs[kDenK][kDenTile + 1];
    const int a0 = blockIdx.x * kDenTile, b0 = blockIdx.y * kDenTile;
    const int k_begin = blockIdx.z * kper, k_end = min(M, k_begin + kper);
    const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = 0.0;
    for (int k0 = k_begin; k0 < k_end; k0 += kDenK) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int e = tid + 256 * q;          // kDenK x kDenTile values per operand
            const int kk = e / kDenTile, r = e % kDenTile;
            const int tr = k0 + kk;
            double va = 0.0, vb = 0.0;
            if (tr < k_end) {
                const T *x = (st[tr].cur ? slot1 : slot0) + (size_t)tr * L;
                va = (double)x[a0 + r] * w[tr];
                vb = (double)x[b0 + r];
            }
            As[kk][r] = va;
            Bs[kk][r] = vb;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kDenK; kk++) {
            double av[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; i++) av[i] = As[kk][ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; j++) bv[j] = Bs[kk][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] += av[i] * bv[j];
        }
        __syncthreads();
    }
    double *out = part + (size_t)blockIdx.z * L * L;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) out[(size_t)(a0 + ty + 16 * i) * L + b0 + tx + 16 * j] = acc[i][j];
}

__global__ void density_reduce_kernel(const double *part, int splits, int d, int dp, int L, double *rho)
{
    const size_t n = (size_t)d * d;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const int a = (int)(e / d), b = (int)(e % d);
        double re = 0.0, im = 0.0;
        for (int s = 0; s < splits; s++) {
            const double *G = part + (size_t)s * L * L;
            re += G[(size_t)a * L + b] + G[(size_t)(dp + a) * L + dp + b];
            im += G[(size_t)(dp + a) * L + b] - G[(size_t)a * L + dp + b];
        }
        rho[2 * e] = re;
        rho[2 * e + 1] = im;
    }
}

// Tr rho^2 = sum |rho_ab|^2 (rho Hermitian): one workgroup, fixed order.
__global__ __launch_bounds__(1024) void purity_kernel(const double *rho, size_t n, double *out)
{
    __shared__ double s[1024];
    double v = 0.0;
    for (size_t e = threadIdx.x; e < n; e += 1024) v += rho[2 * e] * rho[2 * e] + rho[2 * e + 1] * rho[2 * e + 1];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = s[0];
}

void launch_purity(const double *rho, size_t n, double *out, hipStream_t s)
{
    purity_kernel<<<1, 1024, 0, s>>>(rho, n, out);
}

int density_splits(int M, int L)
{
    const int tiles = (L / kDenTile) * (L / kDenTile);
    int sp = std::max(1, std::min(16, 2048 / std::max(1, tiles)));
    sp = std::min(sp, std::max(1, (M + kDenK - 1) / kDenK));
    return sp;
}

void launch_density(int prec, const void *slot0, const void *slot1, const TrajState *st, int M, int d,
                    int dp, int L, double *w, double *part, int splits, double *rho, double *purity,
                    hipStream_t s)
{
    const int wblocks = (M + 3) / 4;          // four waves (trajectories) per workgroup
    int kper = (M + splits - 1) / splits;
    kper = (kper + kDenK - 1) / kDenK * kDenK;
    const dim3 grid(L / kDenTile, L / kDenTile, splits);
    if (prec == kFP64) {
        density_weight_kernel<double><<<wblocks, 256, 0, s>>>((const double *)slot0, (const double *)slot1,
                                                              st, M, L, w);
        density_gram_kernel<double><<<grid, 256, 0, s>>>((const double *)slot0, (const double *)slot1, st, w,
                                                         M, L, kper, part);
    } else {
        density_weight_kernel<float><<<wblocks, 256, 0, s>>>((const float *)slot0, (const float *)slot1, st,
                                                             M, L, w);
        density_gram_kernel<float><<<grid, 256, 0, s>>>((const float *)slot0, (const float *)slot1, st, w, M,
                                                        L, kper, part);
    }
    density_reduce_kernel<<<1024, 256, 0, s>>>(part, splits, d, dp, L, rho);
    purity_kernel<<<1, 1024, 0, s>>>(rho, (size_t)d * d, purity);
}

// ---------------------------------------------------------------------------
// Band table of a packed operand: for each 128-row tile, the K blocks (of 32
// columns, counted within a half of the block-real operand) between the first
// and the last nonzero entry of its rows in either half, and the same for each of
// its four 32-row fragments (one byte each: starts in frag[0], ends in frag[1]).
// The step GEMM skips the blocks outside a tile's range and, inside it, the matrix
// products of a fragment outside the fragment's range: both hold exact zeros only,
// so its results do not change.  The tile range is the union of its fragments'.
// The propagators of a tridiagonal H_eff are banded: the Taylor polynomial of
// A / 2^s has bandwidth 18 and each squaring doubles it, and entries far from the
// diagonal underflow, in FP32 storage soonest.
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline bool nonzero_at(const T *P, const T *Plo, size_t o)
{
    return P[o] != T(0) || (Plo && Plo[o] != T(0));
}

template <typename T>
__global__ __launch_bounds__(256) void band_kernel(const T *P, const T *Plo, int L, int dp, int *out,
                                                   unsigned *frag)
{
    __shared__ int s_lo[4], s_hi[4];
    if (threadIdx.x < 4) {
        s_lo[threadIdx.x] = dp;
        s_hi[threadIdx.x] = 0;
    }
    __syncthreads();
    const int r0 = blockIdx.x * 128;
    for (int f = 0; f < 4; f++) {
        int lo = dp, hi = 0;
        for (int e = threadIdx.x; e < 32 * L; e += blockDim.x) {
            const int r = r0 + 32 * f + e / L, k = e % L;
            if (nonzero_at(P, Plo, (size_t)r * L + k)) {
                const int kk = k < dp ? k : k - dp;
                lo = min(lo, kk);
                hi = max(hi, kk + 1);
            }
        }
        atomicMin(&s_lo[f], lo);
        atomicMax(&s_hi[f], hi);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int lo = dp, hi = 0;
        for (int f = 0; f < 4; f++) {
            if (s_hi[f] > s_lo[f]) {
                lo = min(lo, s_lo[f]);
                hi = max(hi, s_hi[f]);
            }
        }
        const bool any = hi > lo;
        const int kb0 = any ? lo / 32 : 0, kb1 = any ? (hi + 31) / 32 : 0;
        out[2 * blockIdx.x] = kb0;
        out[2 * blockIdx.x + 1] = kb1;
        unsigned w0 = 0, w1 = 0;
        for (int f = 0; f < 4; f++) {
            const bool fa = s_hi[f] > s_lo[f];        // an all-zero fragment gets an empty range
            const unsigned b0 = fa ? (unsigned)(s_lo[f] / 32) : (unsigned)kb0;
            const unsigned b1 = fa ? (unsigned)((s_hi[f] + 31) / 32) : (unsigned)kb0;
            w0 |= b0 << (8 * f);
            w1 |= b1 << (8 * f);
        }
        frag[kFragWords * blockIdx.x] = w0;
        frag[kFragWords * blockIdx.x + 1] = w1;
    }
}

__global__ void band_dense_kernel(int R, int dp, int *out, unsigned *frag)
{
    const int rt = blockIdx.x * blockDim.x + threadIdx.x;
    if (rt >= R) return;
    out[2 * rt] = 0;
    out[2 * rt + 1] = dp / 32;
    frag[kFragWords * rt] = 0;
    frag[kFragWords * rt + 1] = (unsigned)(dp / 32) * 0x01010101u;
}

// Complex-pair layout, 3xBF16: tile rt holds complex rows [64 rt, 64 rt + 64) of P (dp x L),
// fragment f its rows 64 rt + 32 f ..; K blocks count kCpBlock complex columns, and a
// column's Re and Im entries (k and dp + k) fall in the same block.
template <typename T>
__global__ __launch_bounds__(256) void band_cp_kernel(const T *P, const T *Plo, int dp, int *out, unsigned *frag)
{
    __shared__ int s_lo[2], s_hi[2];
    if (threadIdx.x < 2) {
        s_lo[threadIdx.x] = dp;
        s_hi[threadIdx.x] = 0;
    }
    __syncthreads();
    const int L = 2 * dp, r0 = blockIdx.x * kCpRows;
    for (int f = 0; f < 2; f++) {
        int lo = dp, hi = 0;
        for (int e = threadIdx.x; e < 32 * L; e += blockDim.x) {
            const int r = r0 + 32 * f + e / L, k = e % L;
            if (nonzero_at(P, Plo, (size_t)r * L + k)) {
                const int kk = k < dp ? k : k - dp;
                lo = min(lo, kk);
                hi = max(hi, kk + 1);
            }
        }
        atomicMin(&s_lo[f], lo);
        atomicMax(&s_hi[f], hi);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int lo = dp, hi = 0;
        for (int f = 0; f < 2; f++) {
            if (s_hi[f] > s_lo[f]) {
                lo = min(lo, s_lo[f]);
                hi = max(hi, s_hi[f]);
            }
        }
        // an all-zero row tile keeps one block (of exact zeros, with empty fragment
        // ranges), so that every row tile of a work unit has a block to end on
        const bool any = hi > lo;
        const int kb0 = any ? lo / kCpBlock : 0, kb1 = any ? (hi + kCpBlock - 1) / kCpBlock : 1;
        out[2 * blockIdx.x] = kb0;
        out[2 * blockIdx.x + 1] = kb1;
        unsigned w0 = 0, w1 = 0;
        for (int f = 0; f < 2; f++) {
            const bool fa = s_hi[f] > s_lo[f];        // an all-zero fragment gets an empty range
            const unsigned b0 = fa ? (unsigned)(s_lo[f] / kCpBlock) : (unsigned)kb0;
            const unsigned b1 = fa ? (unsigned)((s_hi[f] + kCpBlock - 1) / kCpBlock) : (unsigned)kb0;
            w0 |= b0 << (8 * f);
            w1 |= b1 << (8 * f);
        }
        frag[kFragWords * blockIdx.x] = w0;
        frag[kFragWords * blockIdx.x + 1] = w1;
    }
}

__global__ void band_dense_cp_kernel(int R, int dp, int *out, unsigned *frag)
{
    const int rt = blockIdx.x * blockDim.x + threadIdx.x;
    if (rt >= R) return;
    out[2 * rt] = 0;
    out[2 * rt + 1] = dp / kCpBlock;
    frag[kFragWords * rt] = 0;
    frag[kFragWords * rt + 1] = (unsigned)(dp / kCpBlock) * 0x0101u;
}

// FP32 complex-pair layout: the tile's K blocks as band_cp_kernel, and for each of
// its four 16-row fragments the window [lo, hi) of its nonzero complex columns in
// units of 4 (word f = lo | hi << 16; an all-zero fragment gets an empty window).
__global__ __launch_bounds__(256) void band_cp16_kernel(const float *P, int dp, int *out, unsigned *frag)
{
    __shared__ int s_lo[4], s_hi[4];
    if (threadIdx.x < 4) {
        s_lo[threadIdx.x] = dp;
        s_hi[threadIdx.x] = 0;
    }
    __syncthreads();
    const int L = 2 * dp, r0 = blockIdx.x * kCpRows;
    for (int f = 0; f < 4; f++) {
        int lo = dp, hi = 0;
        for (int e = threadIdx.x; e < 16 * L; e += blockDim.x) {
            const int r = r0 + 16 * f + e / L, k = e % L;
            if (P[(size_t)r * L + k] != 0.f) {
                const int kk = k < dp ? k : k - dp;
                lo = min(lo, kk);
                hi = max(hi, kk + 1);
            }
        }
        atomicMin(&s_lo[f], lo);
        atomicMax(&s_hi[f], hi);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int lo = dp, hi = 0;
        for (int f = 0; f < 4; f++) {
            if (s_hi[f] > s_lo[f]) {
                lo = min(lo, s_lo[f]);
                hi = max(hi, s_hi[f]);
            }
        }
        // an all-zero row tile keeps one block (of exact zeros, with empty windows)
        const bool any = hi > lo;
        const int kb0 = any ? lo / kCpBlock : 0, kb1 = any ? (hi + kCpBlock - 1) / kCpBlock : 1;
        out[2 * blockIdx.x] = kb0;
        out[2 * blockIdx.x + 1] = kb1;
        for (int f = 0; f < 4; f++) {
            const bool fa = s_hi[f] > s_lo[f];
            const unsigned w0 = fa ? (unsigned)(s_lo[f] / 4) : (unsigned)(4 * kb0);
            const unsigned w1 = fa ? (unsigned)((s_hi[f] + 3) / 4) : (unsigned)(4 * kb0);
            frag[kFragWords * blockIdx.x + f] = w0 | (w1 << 16);
        }
    }
}

__global__ void band_dense_cp16_kernel(int R, int dp, int *out, unsigned *frag)
{
    const int rt = blockIdx.x * blockDim.x + threadIdx.x;
    if (rt >= R) return;
    out[2 * rt] = 0;
    out[2 * rt + 1] = dp / kCpBlock;
    for (int f = 0; f < 4; f++) frag[kFragWords * rt + f] = (unsigned)(dp / 4) << 16;
}

void launch_band(int prec, bool cp, const void *P, const void *Plo, int L, int dp, bool dense, int *out,
                 unsigned *frag, hipStream_t s)
{
    if (cp && prec != kBF16x3) {
        const int Rc = dp / kCpRows;
        if (dense) band_dense_cp16_kernel<<<(Rc + 63) / 64, 64, 0, s>>>(Rc, dp, out, frag);
        else band_cp16_kernel<<<Rc, 256, 0, s>>>((const float *)P, dp, out, frag);
        return;
    }
    const int R = L / 128;
    if (cp) {   // 3xBF16
        const int Rc = dp / kCpRows;
        if (dense) band_dense_cp_kernel<<<(Rc + 63) / 64, 64, 0, s>>>(Rc, dp, out, frag);
        else band_cp_kernel<__bf16><<<Rc, 256, 0, s>>>((const __bf16 *)P, (const __bf16 *)Plo, dp, out, frag);
        return;
    }
    if (dense) band_dense_kernel<<<(R + 63) / 64, 64, 0, s>>>(R, dp, out, frag);
    else if (prec == kFP64) band_kernel<double><<<R, 256, 0, s>>>((const double *)P, nullptr, L, dp, out, frag);
    else if (prec == kBF16x3)
        band_kernel<__bf16><<<R, 256, 0, s>>>((const __bf16 *)P, (const __bf16 *)Plo, L, dp, out, frag);
    else band_kernel<float><<<R, 256, 0, s>>>((const float *)P, nullptr, L, dp, out, frag);
}

}  // namespace qjx
