// qjx_kernels.hip -- HIP kernels of the MI355X trajectory engine (gfx950).
//
// One propagation "iteration" advances every active trajectory by exactly one
// matrix-vector step at its own (branch, level) of the bisection tree:
//   step_gemm   : grouped GEMM  Y_g = U_g X_g  over the trajectories listed in
//                 group g = b*D + lvl, columns gathered by trajectory id, output
//                 written to the trajectory's spare slot, FP64 partial ||y||^2
//                 per row tile in the epilogue.
//   step_decide : per trajectory, the FP64 crossing test and the recursion of
//                 one_period_branch (reference qj_proc.cpp:271-343) as a state
//                 machine; jumps (experiment.cpp:1273-1382) are applied by the
//                 whole wavefront; the trajectory is appended to the list of the
//                 group it steps in next.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qjx_internal.h"
#include "qjx_kernels.h"

namespace qjx {

// ---------------------------------------------------------------------------
// MCG59 (MKL VSL_BRNG_MCG59): x <- a x mod 2^59, a = 13^13; output x * 2^-59.
// The per-trajectory leap-frog stream uses the multiplier a^mns.
// ---------------------------------------------------------------------------
__host__ __device__ inline uint64_t mcg_mul(uint64_t x, uint64_t y)
{
    return (x * y) & ((1ULL << 59) - 1ULL);
}

__host__ __device__ inline uint64_t mcg_pow(uint64_t a, uint64_t e)
{
    uint64_t r = 1;
    while (e) {
        if (e & 1ULL) r = mcg_mul(r, a);
        a = mcg_mul(a, a);
        e >>= 1;
    }
    return r;
}

__device__ inline double mcg_next(uint64_t &x, uint64_t stride)
{
    const double u = (double)x * 0x1p-59;
    x = mcg_mul(x, stride);
    return u;
}

// ---------------------------------------------------------------------------
// Small helpers
// ---------------------------------------------------------------------------
template <typename T> struct Real2;
template <> struct Real2<float> { typedef float2 type; };
template <> struct Real2<double> { typedef double2 type; };

__device__ inline double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Prefix over group counts (G <= kMaxGroups), computed by thread 0 into LDS.
__device__ inline void group_prefix(const int *cnt, int G, int per, int *pre)
{
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int g = 0; g < G; g++) {
            pre[g] = acc;
            acc += (cnt[g] + per - 1) / per;
        }
        pre[G] = acc;
    }
    __syncthreads();
}

__device__ inline int find_group(const int *pre, int G, int idx)
{
    int g = 0;
    while (g + 1 < G && pre[g + 1] <= idx) g++;
    return g;
}

// ---------------------------------------------------------------------------
// Grouped propagation GEMM, vector-FMA form (LDS-tiled, 4x4 register tile per
// thread).  Tile BM x BN over (rows of the block-real state, trajectories).
// ---------------------------------------------------------------------------
template <typename T, int BM, int BN, int BK>
__global__ __launch_bounds__(256) void step_gemm_fma(StepArgs a)
{
    __shared__ int s_pre[kMaxGroups + 1];
    __shared__ T As[BK][BM + 4];
    __shared__ T Bs[BK][BN + 1];
    __shared__ int s_tr[BN];
    __shared__ int s_cur[BN];
    __shared__ double s_norm[BM / 4][BN + 1];

    const int tid = threadIdx.x;
    if (blockIdx.x == 0 && blockIdx.y == 0) {
        if (tid < a.G) a.cnt_clear[tid] = 0;
        if (tid == 0) a.jcnt[0] = 0;
    }
    group_prefix(a.cnt, a.G, BN, s_pre);
    const int ct = blockIdx.y;
    if (ct >= s_pre[a.G]) return;
    const int g = find_group(s_pre, a.G, ct);
    const int col0 = (ct - s_pre[g]) * BN;
    const int ncol = min(BN, a.cnt[g] - col0);
    const size_t LL = (size_t)a.L * a.L;
    const T *U = reinterpret_cast<const T *>(a.props) + (size_t)g * LL;
    if (tid < BN) {
        if (tid < ncol) {
            const int tr = a.list[(size_t)g * a.M + col0 + tid];
            s_tr[tid] = tr;
            s_cur[tid] = a.st[tr].cur;
        } else {
            s_tr[tid] = -1;
            s_cur[tid] = 0;
        }
    }
    __syncthreads();

    const int r0 = blockIdx.x * BM;
    const int tx = tid % 16, ty = tid / 16;
    T acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = T(0);

    // load mapping: 4 consecutive k of one row of U (row-major) / one column of X
    const int a_i = tid / (BK / 4), a_k = (tid % (BK / 4)) * 4;      // BM x BK / 4 threads
    const int b_j = tid / (BK / 4), b_k = (tid % (BK / 4)) * 4;      // BN x BK / 4 threads
    const T *xcol = nullptr;
    if (b_j < BN && s_tr[b_j] >= 0)
        xcol = reinterpret_cast<const T *>(a.slot[s_cur[b_j]]) + (size_t)s_tr[b_j] * a.L;

    for (int k0 = 0; k0 < a.L; k0 += BK) {
        if (a_i < BM) {
            const T *src = U + (size_t)(r0 + a_i) * a.L + k0 + a_k;
#pragma unroll
            for (int q = 0; q < 4; q++) As[a_k + q][a_i] = src[q];
        }
        if (b_j < BN) {
#pragma unroll
            for (int q = 0; q < 4; q++) Bs[b_k + q][b_j] = xcol ? xcol[k0 + b_k + q] : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; kk++) {
            T av[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; i++) av[i] = As[kk][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; j++) bv[j] = Bs[kk][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] += av[i] * bv[j];
        }
        __syncthreads();
    }

#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int c = tx + 16 * j;
        double ns = 0.0;
        if (c < ncol) {
            T *y = reinterpret_cast<T *>(a.slot[s_cur[c] ^ 1]) + (size_t)s_tr[c] * a.L + r0 + ty * 4;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                y[i] = acc[i][j];
                const double v = (double)acc[i][j];
                ns += v * v;
            }
        }
        s_norm[ty][c] = ns;
    }
    __syncthreads();
    if (tid < ncol) {
        double s = 0.0;
        for (int q = 0; q < BM / 4; q++) s += s_norm[q][tid];
        a.normpart[(size_t)blockIdx.x * a.M + s_tr[tid]] = s;
    }
}

// ---------------------------------------------------------------------------
// Decide kernel: crossing test, recursion bookkeeping, jumps, next group lists.
// ---------------------------------------------------------------------------
// After a completed step at level l (accepted, or the finest-level step that
// carried a jump; t has already advanced by dt_l): counter++ and, while levels
// finish, return to the parent (qj_proc.cpp:325-342: t -= dt_parent, then
// counter++ and t += dt_parent).  Level 0 finishes after S/2 steps and closes
// the half period; a level l >= 1 finishes after 2 steps.
__device__ inline void complete_step(TrajState &s, int l, const double *dt, int S)
{
    for (;;) {
        if (l == 0) {
            s.cnt0++;
            s.pos++;
            if (s.cnt0 == S / 2) {
                s.cnt0 = 0;
                s.b ^= 1;
            }
            break;
        }
        const uint32_t bit = 1u << l;
        if (!(s.bits & bit)) {
            s.bits |= bit;
            break;
        }
        s.bits &= ~bit;
        l -= 1;
        s.t -= dt[l];
        s.t += dt[l];
    }
    s.lvl = l;
}

template <typename T>
__global__ __launch_bounds__(256) void step_decide(StepArgs a, uint64_t stride)
{
    __shared__ int s_pre[kMaxGroups + 1];
    __shared__ int s_cnt[kMaxGroups + 1];    // per-group appends of this block (+ jumps)
    __shared__ int s_base[kMaxGroups + 1];
    if (threadIdx.x <= kMaxGroups) s_cnt[threadIdx.x] = 0;
    group_prefix(a.cnt, a.G, 1, s_pre);      // (ends with a barrier)
    const int total = s_pre[a.G];
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx == 0) atomicAdd(&a.counters[0], (unsigned long long)total);
    if ((int)(blockIdx.x * blockDim.x) >= total) return;   // whole block idle
    // appends are reserved per block: LDS counters first, one global atomic per group
    int app_g = -1, app_local = 0;           // next-group append
    int jmp_local = -1;                      // jump-list append
    int tr = -1, jsrc = 0;
    double n = 0.0;
    if (idx < total) {
        const int g = find_group(s_pre, a.G, idx);
        tr = a.list[(size_t)g * a.M + (idx - s_pre[g])];
        TrajState s = a.st[tr];
        for (int r = 0; r < a.R; r++) n += a.normpart[(size_t)r * a.M + tr];
        const int lvl = s.lvl;
        const double dt = a.dt[lvl];
        bool alive = true;
        if (n > 1.0 + a.norm_tol) {
            // experiment.cpp:1247-1251 throws: record the first offender, stop it.
            if (atomicCAS(&a.err[0], 0, 1) == 0) {
                a.err[1] = tr;
                const unsigned long long bits = (unsigned long long)__double_as_longlong(n);
                a.err[2] = (int)(bits & 0xffffffffULL);
                a.err[3] = (int)(bits >> 32);
            }
            alive = false;
        } else {
            const bool crossed = !(n > s.eta);             // experiment.cpp:1253-1260
            if (lvl == a.D - 1) {
                if (crossed) {
                    if (a.jump > 0 && a.is_obs == 1) {
                        if (s.logn < a.log.cap) {
                            const size_t o = (size_t)tr * a.log.cap + s.logn;
                            a.log.time[o] = s.t + dt;
                            a.log.norm[o] = n;
                            a.log.eta[o] = s.eta;
                            a.log.type[o] = 0;
                        }
                        s.logn++;
                        s.njumps++;
                    }
                    // recovery draws `ran` (one channel, so channel 0 is chosen),
                    // then eta is redrawn while it is 0 (qj_proc.cpp:310-314).
                    (void)mcg_next(s.rng, stride);
                    double e = mcg_next(s.rng, stride);
                    while (e == 0.0) e = mcg_next(s.rng, stride);
                    s.eta = e;
                    // the jump kernel reads the step just taken (slot cur ^ 1) and
                    // writes the jumped state over the pre-step state (slot cur)
                    jmp_local = atomicAdd(&s_cnt[kMaxGroups], 1);
                    jsrc = s.cur ^ 1;
                } else {
                    s.cur ^= 1;
                }
                s.t += dt;
                complete_step(s, lvl, a.dt, a.S);
            } else if (crossed) {
                s.lvl = lvl + 1;          // bisect from the pre-step state
            } else {
                s.cur ^= 1;
                s.t += dt;
                complete_step(s, lvl, a.dt, a.S);
            }
        }
        a.st[tr] = s;
        if (alive && s.pos < s.target) {
            app_g = s.b * a.D + s.lvl;
            app_local = atomicAdd(&s_cnt[app_g], 1);
        }
    }
    __syncthreads();
    if (threadIdx.x < a.G) {
        const int c = s_cnt[threadIdx.x];
        s_base[threadIdx.x] = c ? atomicAdd(&a.cnt_next[threadIdx.x], c) : 0;
    } else if (threadIdx.x == kMaxGroups) {
        const int c = s_cnt[kMaxGroups];
        s_base[kMaxGroups] = c ? atomicAdd(a.jcnt, c) : 0;
    }
    __syncthreads();
    if (app_g >= 0) a.list_next[(size_t)app_g * a.M + s_base[app_g] + app_local] = tr;
    if (jmp_local >= 0) {
        const int j = s_base[kMaxGroups] + jmp_local;
        a.jlist[j] = tr;
        a.jsrc[j] = jsrc;
        a.jnorm[j] = n;
    }
}

// ---------------------------------------------------------------------------
// Jumps: recovery (experiment.cpp:1273-1382) for the single dimer dissipator,
// one workgroup per jumping trajectory (persistent over the jump list):
//   phi <- phi / ||phi||;  v = L phi;  phi <- v / sqrt(||v||^2)
// in FP64.  L is tridiagonal: L_kk = Ldiag[k], L_{k+1,k} = -Loff[k],
// L_{k,k+1} = +Loff[k]; v_k sums its three terms in ascending column order.
// ---------------------------------------------------------------------------
constexpr int kJumpThreads = 256;
constexpr int kMaxDim = 2048;                       // largest supported d
constexpr int kJumpPer = kMaxDim / kJumpThreads;    // elements per thread

template <typename T>
__global__ __launch_bounds__(kJumpThreads) void step_jump(StepArgs a)
{
    __shared__ double pre[kMaxDim + 2];
    __shared__ double pim[kMaxDim + 2];
    __shared__ double red[kJumpThreads / 64];
    const int njump = a.jcnt[0];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int d = a.d, dp = a.dp;
    if (blockIdx.x == 0 && tid == 0 && njump > 0) atomicAdd(&a.counters[1], (unsigned long long)njump);
    for (int j = blockIdx.x; j < njump; j += gridDim.x) {
        const int tr = a.jlist[j];
        const int src = a.jsrc[j];
        const double sn = sqrt(a.jnorm[j]);
        const T *x = reinterpret_cast<const T *>(a.slot[src]) + (size_t)tr * a.L;
        T *y = reinterpret_cast<T *>(a.slot[src ^ 1]) + (size_t)tr * a.L;
        // pre/pim[k + 1] = phi_k / ||phi||, zero-padded at both ends
        for (int k = tid; k < d; k += kJumpThreads) {
            pre[k + 1] = (double)x[k] / sn;
            pim[k + 1] = (double)x[dp + k] / sn;
        }
        if (tid == 0) {
            pre[0] = pim[0] = 0.0;
            pre[d + 1] = pim[d + 1] = 0.0;
        }
        __syncthreads();
        double vre[kJumpPer], vim[kJumpPer];
        double p = 0.0;
#pragma unroll
        for (int m = 0; m < kJumpPer; m++) {
            const int k = tid + m * kJumpThreads;
            vre[m] = vim[m] = 0.0;
            if (k < d) {
                double r = 0.0, i = 0.0;
                if (k > 0) {
                    const double l = -a.Loff[k - 1];
                    r += l * pre[k];
                    i += l * pim[k];
                }
                r += a.Ldiag[k] * pre[k + 1];
                i += a.Ldiag[k] * pim[k + 1];
                if (k + 1 < d) {
                    const double l = a.Loff[k];
                    r += l * pre[k + 2];
                    i += l * pim[k + 2];
                }
                vre[m] = r;
                vim[m] = i;
                p += r * r + i * i;
            }
        }
        p = wave_sum(p);
        if (lane == 0) red[wid] = p;
        __syncthreads();
        double tot = 0.0;
#pragma unroll
        for (int w = 0; w < kJumpThreads / 64; w++) tot += red[w];
        const double sp = sqrt(tot / 1.0);
#pragma unroll
        for (int m = 0; m < kJumpPer; m++) {
            const int k = tid + m * kJumpThreads;
            if (k < d) {
                y[k] = (T)(vre[m] / sp);
                y[dp + k] = (T)(vim[m] / sp);
            }
        }
        __syncthreads();   // LDS reuse by the next jump of this workgroup
    }
}

}  // namespace qjx

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

__global__ void zero_ints_kernel(int *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Every trajectory sits at a top-level step boundary between runs, so it starts
// at level 0 of its current branch.
__global__ void seed_lists_kernel(TrajState *st, int M, int n, int *list, int *cnt, int D)
{
    const int tr = blockIdx.x * blockDim.x + threadIdx.x;
    if (tr >= M) return;
    TrajState s = st[tr];
    s.target = s.pos + n;
    st[tr] = s;
    if (n > 0) {
        const int g = s.b * D;
        const int p = atomicAdd(&cnt[g], 1);
        list[(size_t)g * M + p] = tr;
    }
}

void launch_seed_lists(TrajState *st, int M, int n_steps, int *list, int *cnt, int G,
                       hipStream_t s)
{
    zero_ints_kernel<<<1, kMaxGroups, 0, s>>>(cnt, kMaxGroups);
    seed_lists_kernel<<<(M + 255) / 256, 256, 0, s>>>(st, M, n_steps, list, cnt, G / 2);
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
// Propagation iteration
// ---------------------------------------------------------------------------
void launch_step_gemm(int prec, int gemm_kind, const StepArgs &a, hipStream_t s)
{
    if (gemm_kind == kGemmMfma) {
        launch_step_gemm_mfma(prec, a, s);
        return;
    }
    constexpr int BM = 64, BN = 64, BK = 16;
    dim3 grid(a.R, (a.M + BN - 1) / BN + a.G);
    if (prec == kFP64) step_gemm_fma<double, BM, BN, BK><<<grid, 256, 0, s>>>(a);
    else step_gemm_fma<float, BM, BN, BK><<<grid, 256, 0, s>>>(a);
}

void launch_step_decide(int prec, const StepArgs &a, uint64_t stride, hipStream_t s)
{
    const int jblocks = a.M < 2048 ? a.M : 2048;
    if (prec == kFP64) {
        step_decide<double><<<(a.M + 255) / 256, 256, 0, s>>>(a, stride);
        step_jump<double><<<jblocks, kJumpThreads, 0, s>>>(a);
    } else {
        step_decide<float><<<(a.M + 255) / 256, 256, 0, s>>>(a, stride);
        step_jump<float><<<jblocks, kJumpThreads, 0, s>>>(a);
    }
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

void launch_pack_propagator(int prec, int n, int dp, const double *U, void *dst, void *dst_lo,
                            hipStream_t s)
{
    if (prec == kFP64) pack_propagator_kernel<double><<<2048, 256, 0, s>>>(n, dp, U, (double *)dst);
    else if (prec == kBF16x3)
        pack_propagator_bf16x3_kernel<<<2048, 256, 0, s>>>(n, dp, U, (__bf16 *)dst, (__bf16 *)dst_lo);
    else pack_propagator_kernel<float><<<2048, 256, 0, s>>>(n, dp, U, (float *)dst);
}

}  // namespace qjx
