// qjx_gemm_mfma.hip -- grouped propagation GEMM on CDNA4 matrix cores.
//
//   Y[:, j] = U_g X[:, j]   for every trajectory j listed in group g = b*D + lvl
//
// U_g is the block-real propagator [[Re U, -Im U], [Im U, Re U]] (L x L, L = 2 dp),
// stored row-major for these kernels so that a row segment of the A tile and a
// column segment of the X tile are both contiguous 16-byte loads straight into
// LDS ([row][k] and [col][k] layouts).  Inside a K block of 32 the MFMA k-steps
// visit k in a permuted order (lane half / quarter h takes k = s + (32/H) h),
// which lets every lane read its A and B fragments for the whole K block with
// 16-byte LDS reads; the permutation is the same for every tile and column, so
// each output element's accumulation order is independent of where its column
// lands (results do not depend on batch composition).
//
// FP32 form : v_mfma_f32_32x32x2f32, 4 waves as 2 x 2, wave tile 64 x 64, WG tile 128 x 128.
// FP64 form : v_mfma_f64_16x16x4f64, 8 waves as 4 x 2, wave tile 32 x 64, WG tile 128 x 128.
// BF16x3    : v_mfma_f32_32x32x16_bf16 x 3 per block (split operands), layout as FP32.
// Epilogue  : output to the trajectory's spare slot, FP64 sum of squares per column over
//             the WG's 128 rows into normpart[row tile][trajectory].
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qjx_internal.h"
#include "qjx_kernels.h"

namespace qjx {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

constexpr int MF_BM = 128, MF_BN = 128, MF_BK = 32;
constexpr int MF_THREADS = 256;

__device__ inline void mf_group_prefix(const int *cnt, int G, int per, int *pre)
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

__device__ inline int mf_find_group(const int *pre, int G, int idx)
{
    int g = 0;
    while (g + 1 < G && pre[g + 1] <= idx) g++;
    return g;
}

// Resolve this workgroup's tile: group, column range, trajectory ids and slots.
// Returns false if the block has no tile this iteration.
__device__ inline bool mf_tile(const StepArgs &a, int *s_pre, int *s_tr, int *s_cur, int &g,
                               int &ncol)
{
    const int tid = threadIdx.x;
    if (blockIdx.x == 0 && blockIdx.y == 0) {
        if (tid < a.G) a.cnt_clear[tid] = 0;
        if (tid == 0) a.jcnt[0] = 0;
    }
    mf_group_prefix(a.cnt, a.G, MF_BN, s_pre);
    const int ct = blockIdx.y;
    if (ct >= s_pre[a.G]) return false;
    g = mf_find_group(s_pre, a.G, ct);
    const int col0 = (ct - s_pre[g]) * MF_BN;
    ncol = min(MF_BN, a.cnt[g] - col0);
    if (tid < MF_BN) {
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
    return true;
}

// ---------------------------------------------------------------------------
// FP32: v_mfma_f32_32x32x2f32
//   A (32 x 2): lane l holds A[l % 32][l / 32]
//   B (2 x 32): lane l holds B[l / 32][l % 32]
//   D (32 x 32): lane l, register r holds D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32]
// Within a K block of 32, MFMA step s (0..15) uses k = s + 16 h for lane half h.
// ---------------------------------------------------------------------------
constexpr int F32_PITCH = MF_BK + 4;   // LDS row pitch in floats (16-byte aligned, skewed banks)

__global__ __launch_bounds__(MF_THREADS) void step_gemm_mfma_f32(StepArgs a)
{
    __shared__ int s_pre[kMaxGroups + 1];
    __shared__ int s_tr[MF_BN];
    __shared__ int s_cur[MF_BN];
    __shared__ __attribute__((aligned(16))) float As[MF_BM * F32_PITCH];
    __shared__ __attribute__((aligned(16))) float Bs[MF_BN * F32_PITCH];
    __shared__ double s_norm[2][MF_BN];

    int g, ncol;
    if (!mf_tile(a, s_pre, s_tr, s_cur, g, ncol)) return;

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;             // wave position in the 2 x 2 layout
    const int L = a.L;
    const int r0 = blockIdx.x * MF_BM;
    const float *U = reinterpret_cast<const float *>(a.props) + (size_t)g * L * L;

    // global -> LDS copy mapping: 128 rows (or columns) x 32 k = 1024 float4,
    // 4 per thread: row = tid / 8 + 32 q, k4 = (tid % 8) * 4
    const int cp_k = (tid & 7) * 4;
    const int cp_r = tid >> 3;
    const float *arow[4];
    const float *bcol[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int r = cp_r + 32 * q;
        arow[q] = U + (size_t)(r0 + r) * L + cp_k;
        const int cc = r < ncol ? r : 0;     // columns past ncol re-read a valid one (never stored)
        bcol[q] = reinterpret_cast<const float *>(a.slot[s_cur[cc]]) + (size_t)s_tr[cc] * L + cp_k;
    }

    floatx16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[i][j][e] = 0.f;

    floatx4 ra[4], rb[4];
    auto gload = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            ra[q] = *reinterpret_cast<const floatx4 *>(arow[q] + k0);
            rb[q] = *reinterpret_cast<const floatx4 *>(bcol[q] + k0);
        }
    };
    auto lstore = [&]() {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int r = cp_r + 32 * q;
            *reinterpret_cast<floatx4 *>(&As[r * F32_PITCH + cp_k]) = ra[q];
            *reinterpret_cast<floatx4 *>(&Bs[r * F32_PITCH + cp_k]) = rb[q];
        }
    };

    const int h = lane >> 5, li = lane & 31;
    gload(0);
    lstore();
    __syncthreads();
    for (int k0 = 0; k0 < L; k0 += MF_BK) {
        const bool more = k0 + MF_BK < L;
        if (more) gload(k0 + MF_BK);
        // fragments: 16 consecutive k (= s + 16 h) per lane, 4 x 16-byte reads each
        floatx4 fa[2][4], fb[2][4];
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const float *pa = &As[(wm * 64 + i * 32 + li) * F32_PITCH + 16 * h];
            const float *pb = &Bs[(wn * 64 + i * 32 + li) * F32_PITCH + 16 * h];
#pragma unroll
            for (int v = 0; v < 4; v++) {
                fa[i][v] = *reinterpret_cast<const floatx4 *>(pa + 4 * v);
                fb[i][v] = *reinterpret_cast<const floatx4 *>(pb + 4 * v);
            }
        }
#pragma unroll
        for (int s = 0; s < 16; s++) {
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[i][s >> 2][s & 3], fb[j][s >> 2][s & 3],
                                                                      acc[i][j], 0, 0, 0);
        }
        __syncthreads();
        if (more) {
            lstore();
            __syncthreads();
        }
    }

    // epilogue: rows 8 (r/4) + 4 h + (r%4) of each 32 x 32 block, column li
    double ns[2] = {0.0, 0.0};
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int c = wn * 64 + j * 32 + li;
        const bool ok = c < ncol;
        float *y = reinterpret_cast<float *>(a.slot[s_cur[ok ? c : 0] ^ 1]) + (size_t)s_tr[ok ? c : 0] * L;
#pragma unroll
        for (int i = 0; i < 2; i++) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int row = r0 + wm * 64 + i * 32 + 8 * q + 4 * h;
                floatx4 v = {acc[i][j][4 * q], acc[i][j][4 * q + 1], acc[i][j][4 * q + 2], acc[i][j][4 * q + 3]};
                if (ok) *reinterpret_cast<floatx4 *>(y + row) = v;
#pragma unroll
                for (int e = 0; e < 4; e++) ns[j] += (double)v[e] * (double)v[e];
            }
        }
    }
#pragma unroll
    for (int j = 0; j < 2; j++) ns[j] += __shfl_xor(ns[j], 32, 64);
    if (h == 0) {
#pragma unroll
        for (int j = 0; j < 2; j++) s_norm[wm][wn * 64 + j * 32 + li] = ns[j];
    }
    __syncthreads();
    if (tid < ncol) a.normpart[(size_t)blockIdx.x * a.M + s_tr[tid]] = s_norm[0][tid] + s_norm[1][tid];
}

// ---------------------------------------------------------------------------
// FP64: v_mfma_f64_16x16x4f64
//   A (16 x 4): lane l holds A[l % 16][l / 16]
//   B (4 x 16): lane l holds B[l / 16][l % 16]
//   D (16 x 16): lane l, register r holds D[(l / 16) + 4 r][l % 16]
// 8 waves as 4 (rows) x 2 (columns), wave tile 32 x 64 (2 x 4 blocks), WG tile 128 x 128.
// Within a K block of 32, MFMA step s (0..7) uses k = s + 8 q for lane quarter q.
// ---------------------------------------------------------------------------
constexpr int F64_PITCH = MF_BK + 2;   // doubles (16-byte aligned rows, skewed banks)
constexpr int F64_THREADS = 512;

__global__ __launch_bounds__(F64_THREADS) void step_gemm_mfma_f64(StepArgs a)
{
    typedef double double2v __attribute__((ext_vector_type(2)));
    __shared__ int s_pre[kMaxGroups + 1];
    __shared__ int s_tr[MF_BN];
    __shared__ int s_cur[MF_BN];
    __shared__ __attribute__((aligned(16))) double As[MF_BM * F64_PITCH];
    __shared__ __attribute__((aligned(16))) double Bs[MF_BN * F64_PITCH];
    __shared__ double s_norm[4][MF_BN];

    int g, ncol;
    if (!mf_tile(a, s_pre, s_tr, s_cur, g, ncol)) return;

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;             // 4 x 2 waves
    const int L = a.L;
    const int r0 = blockIdx.x * MF_BM;
    const double *U = reinterpret_cast<const double *>(a.props) + (size_t)g * L * L;

    // global -> LDS: 128 rows x 32 k = 2048 double2, 4 per thread:
    // row = tid / 16 + 32 q, k2 = (tid % 16) * 2.  Columns past ncol re-read a
    // valid column (their results are never stored).
    const int cp_k = (tid & 15) * 2;
    const int cp_r = tid >> 4;
    const double *arow = U + (size_t)(r0 + cp_r) * L + cp_k;
    const size_t astep = (size_t)32 * L;
    const double *bcol[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int c = cp_r + 32 * q;
        const int cc = c < ncol ? c : 0;
        bcol[q] = reinterpret_cast<const double *>(a.slot[s_cur[cc]]) + (size_t)s_tr[cc] * L + cp_k;
    }

    doublex4 acc[2][4];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = doublex4{0.0, 0.0, 0.0, 0.0};

    double2v ra[4], rb[4];
    auto gload = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            ra[q] = *reinterpret_cast<const double2v *>(arow + q * astep + k0);
            rb[q] = *reinterpret_cast<const double2v *>(bcol[q] + k0);
        }
    };
    auto lstore = [&]() {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int r = cp_r + 32 * q;
            *reinterpret_cast<double2v *>(&As[r * F64_PITCH + cp_k]) = ra[q];
            *reinterpret_cast<double2v *>(&Bs[r * F64_PITCH + cp_k]) = rb[q];
        }
    };

    const int qq = lane >> 4, li = lane & 15;
    const double *pa0 = &As[(wm * 32 + li) * F64_PITCH + 8 * qq];
    const double *pb0 = &Bs[(wn * 64 + li) * F64_PITCH + 8 * qq];
    gload(0);
    lstore();
    __syncthreads();
    for (int k0 = 0; k0 < L; k0 += MF_BK) {
        const bool more = k0 + MF_BK < L;
        if (more) gload(k0 + MF_BK);
#pragma unroll
        for (int half = 0; half < 2; half++) {
            // 4 consecutive k (= s + 8 qq, s = 4 half .. 4 half + 3) per lane and block
            double2v fa[2][2], fb[4][2];
#pragma unroll
            for (int i = 0; i < 2; i++) {
                const double *pa = pa0 + i * 16 * F64_PITCH + 4 * half;
                fa[i][0] = *reinterpret_cast<const double2v *>(pa);
                fa[i][1] = *reinterpret_cast<const double2v *>(pa + 2);
            }
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const double *pb = pb0 + j * 16 * F64_PITCH + 4 * half;
                fb[j][0] = *reinterpret_cast<const double2v *>(pb);
                fb[j][1] = *reinterpret_cast<const double2v *>(pb + 2);
            }
#pragma unroll
            for (int s4 = 0; s4 < 4; s4++)
#pragma unroll
                for (int i = 0; i < 2; i++)
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i][s4 >> 1][s4 & 1],
                                                                          fb[j][s4 >> 1][s4 & 1], acc[i][j], 0,
                                                                          0, 0);
        }
        __syncthreads();
        if (more) {
            lstore();
            __syncthreads();
        }
    }

    double ns[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int c = wn * 64 + j * 16 + li;
        const bool ok = c < ncol;
        double *y = reinterpret_cast<double *>(a.slot[s_cur[ok ? c : 0] ^ 1]) + (size_t)s_tr[ok ? c : 0] * L;
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const int row = r0 + wm * 32 + i * 16 + qq;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                if (ok) y[row + 4 * e] = acc[i][j][e];
                ns[j] += acc[i][j][e] * acc[i][j][e];
            }
        }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        ns[j] += __shfl_xor(ns[j], 16, 64);
        ns[j] += __shfl_xor(ns[j], 32, 64);
    }
    if (qq == 0) {
#pragma unroll
        for (int j = 0; j < 4; j++) s_norm[wm][wn * 64 + j * 16 + li] = ns[j];
    }
    __syncthreads();
    if (tid < ncol)
        a.normpart[(size_t)blockIdx.x * a.M + s_tr[tid]] =
            ((s_norm[0][tid] + s_norm[1][tid]) + s_norm[2][tid]) + s_norm[3][tid];
}

// ---------------------------------------------------------------------------
// 3 x BF16 split product (FP32 accumulate): with x = x_hi + x_lo (x_hi = bf16(x),
// x_lo = bf16(x - x_hi)) for both operands, U X ~= Uhi Xhi + Uhi Xlo + Ulo Xhi; the
// dropped Ulo Xlo term and the split residue are O(2^-16) relative.
// v_mfma_f32_32x32x16_bf16: A (32 x 16) lane l holds 8 consecutive k of row l % 32,
// B (16 x 32) the same 8 k of column l % 32 (lane half h = l / 32 selects which 8),
// D as in the FP32 form.  The propagator halves are split once when packed; the
// state is split when its tile is staged into LDS.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
constexpr int BF_PITCH = MF_BK + 8;    // bf16 per LDS row (80 bytes: 16-byte aligned, skewed)

__global__ __launch_bounds__(MF_THREADS) void step_gemm_mfma_bf16x3(StepArgs a)
{
    __shared__ int s_pre[kMaxGroups + 1];
    __shared__ int s_tr[MF_BN];
    __shared__ int s_cur[MF_BN];
    __shared__ __attribute__((aligned(16))) __bf16 Ah[MF_BM * BF_PITCH];
    __shared__ __attribute__((aligned(16))) __bf16 Al[MF_BM * BF_PITCH];
    __shared__ __attribute__((aligned(16))) __bf16 Bh[MF_BN * BF_PITCH];
    __shared__ __attribute__((aligned(16))) __bf16 Bl[MF_BN * BF_PITCH];
    __shared__ double s_norm[2][MF_BN];

    int g, ncol;
    if (!mf_tile(a, s_pre, s_tr, s_cur, g, ncol)) return;

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    const int L = a.L;
    const int r0 = blockIdx.x * MF_BM;
    const size_t LL = (size_t)L * L;
    const __bf16 *Uh = reinterpret_cast<const __bf16 *>(a.props) + (size_t)g * LL;
    const __bf16 *Ul = reinterpret_cast<const __bf16 *>(a.props_lo) + (size_t)g * LL;

    // A tiles: 128 rows x 32 k bf16 = 512 chunks of 8, 2 per thread per half:
    //   row = tid / 4 + 64 q, k8 = (tid % 4) * 8
    const int ca_k = (tid & 3) * 8;
    const int ca_r = tid >> 2;
    // B tile: 128 columns x 32 k floats = 1024 float4, 4 per thread:
    //   column = tid / 8 + 32 q, k4 = (tid % 8) * 4
    const int cb_k = (tid & 7) * 4;
    const int cb_r = tid >> 3;
    const float *bcol[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int c = cb_r + 32 * q;
        const int cc = c < ncol ? c : 0;
        bcol[q] = reinterpret_cast<const float *>(a.slot[s_cur[cc]]) + (size_t)s_tr[cc] * L + cb_k;
    }
    floatx16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[i][j][e] = 0.f;

    bf16x8 rah[2], ral[2];
    floatx4 rb[4];
    auto gload = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const size_t o = (size_t)(r0 + ca_r + 64 * q) * L + k0 + ca_k;
            rah[q] = *reinterpret_cast<const bf16x8 *>(Uh + o);
            ral[q] = *reinterpret_cast<const bf16x8 *>(Ul + o);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) rb[q] = *reinterpret_cast<const floatx4 *>(bcol[q] + k0);
    };
    auto lstore = [&]() {
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int r = ca_r + 64 * q;
            *reinterpret_cast<bf16x8 *>(&Ah[r * BF_PITCH + ca_k]) = rah[q];
            *reinterpret_cast<bf16x8 *>(&Al[r * BF_PITCH + ca_k]) = ral[q];
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int r = cb_r + 32 * q;
            bf16x4 h4, l4;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const float x = rb[q][e];
                const __bf16 hb = (__bf16)x;
                h4[e] = hb;
                l4[e] = (__bf16)(x - (float)hb);
            }
            *reinterpret_cast<bf16x4 *>(&Bh[r * BF_PITCH + cb_k]) = h4;
            *reinterpret_cast<bf16x4 *>(&Bl[r * BF_PITCH + cb_k]) = l4;
        }
    };

    const int h = lane >> 5, li = lane & 31;
    gload(0);
    lstore();
    __syncthreads();
    for (int k0 = 0; k0 < L; k0 += MF_BK) {
        const bool more = k0 + MF_BK < L;
        if (more) gload(k0 + MF_BK);
#pragma unroll
        for (int c = 0; c < MF_BK / 16; c++) {
            bf16x8 fah[2], fal[2], fbh[2], fbl[2];
#pragma unroll
            for (int i = 0; i < 2; i++) {
                const int ra = (wm * 64 + i * 32 + li) * BF_PITCH + 16 * c + 8 * h;
                const int rbo = (wn * 64 + i * 32 + li) * BF_PITCH + 16 * c + 8 * h;
                fah[i] = *reinterpret_cast<const bf16x8 *>(&Ah[ra]);
                fal[i] = *reinterpret_cast<const bf16x8 *>(&Al[ra]);
                fbh[i] = *reinterpret_cast<const bf16x8 *>(&Bh[rbo]);
                fbl[i] = *reinterpret_cast<const bf16x8 *>(&Bl[rbo]);
            }
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fal[i], fbh[j], acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fah[i], fbl[j], acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fah[i], fbh[j], acc[i][j], 0, 0, 0);
                }
        }
        __syncthreads();
        if (more) {
            lstore();
            __syncthreads();
        }
    }

    double ns[2] = {0.0, 0.0};
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int c = wn * 64 + j * 32 + li;
        const bool ok = c < ncol;
        float *y = reinterpret_cast<float *>(a.slot[s_cur[ok ? c : 0] ^ 1]) + (size_t)s_tr[ok ? c : 0] * L;
#pragma unroll
        for (int i = 0; i < 2; i++) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int row = r0 + wm * 64 + i * 32 + 8 * q + 4 * h;
                floatx4 v = {acc[i][j][4 * q], acc[i][j][4 * q + 1], acc[i][j][4 * q + 2], acc[i][j][4 * q + 3]};
                if (ok) *reinterpret_cast<floatx4 *>(y + row) = v;
#pragma unroll
                for (int e = 0; e < 4; e++) ns[j] += (double)v[e] * (double)v[e];
            }
        }
    }
#pragma unroll
    for (int j = 0; j < 2; j++) ns[j] += __shfl_xor(ns[j], 32, 64);
    if (h == 0) {
#pragma unroll
        for (int j = 0; j < 2; j++) s_norm[wm][wn * 64 + j * 32 + li] = ns[j];
    }
    __syncthreads();
    if (tid < ncol) a.normpart[(size_t)blockIdx.x * a.M + s_tr[tid]] = s_norm[0][tid] + s_norm[1][tid];
}

void launch_step_gemm_mfma(int prec, const StepArgs &a, hipStream_t s)
{
    dim3 grid(a.L / MF_BM, (a.M + MF_BN - 1) / MF_BN + a.G);
    if (prec == kFP64) step_gemm_mfma_f64<<<grid, F64_THREADS, 0, s>>>(a);
    else if (prec == kBF16x3) step_gemm_mfma_bf16x3<<<grid, MF_THREADS, 0, s>>>(a);
    else step_gemm_mfma_f32<<<grid, MF_THREADS, 0, s>>>(a);
}

// ---------------------------------------------------------------------------
// Lane-layout probe for the two MFMA builtins: one wave computes D = A B for a
// single block with the operand/result layouts documented above and writes D
// row-major.  Used by a test against a host product.
// ---------------------------------------------------------------------------
__global__ void mfma_probe_kernel(const float *A32, const float *B32, float *D32, const double *A64,
                                  const double *B64, double *D64, const float *Abf, const float *Bbf,
                                  float *Dbf)
{
    const int l = threadIdx.x;
    {
        // bf16 32x32x16: A 32 x 16, B 16 x 32 (row-major inputs, exactly representable)
        bf16x8 av, bv;
        for (int e = 0; e < 8; e++) {
            av[e] = (__bf16)Abf[(l % 32) * 16 + 8 * (l / 32) + e];
            bv[e] = (__bf16)Bbf[(8 * (l / 32) + e) * 32 + l % 32];
        }
        floatx16 cb = {};
        cb = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, cb, 0, 0, 0);
        for (int r = 0; r < 16; r++) Dbf[(8 * (r / 4) + 4 * (l / 32) + r % 4) * 32 + l % 32] = cb[r];
    }
    floatx16 c = {};
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(A32[(l % 32) * 2 + l / 32], B32[(l / 32) * 32 + l % 32], c, 0, 0, 0);
    for (int r = 0; r < 16; r++) D32[(8 * (r / 4) + 4 * (l / 32) + r % 4) * 32 + l % 32] = c[r];
    doublex4 e = {0.0, 0.0, 0.0, 0.0};
    e = __builtin_amdgcn_mfma_f64_16x16x4f64(A64[(l % 16) * 4 + l / 16], B64[(l / 16) * 16 + l % 16], e, 0, 0, 0);
    for (int r = 0; r < 4; r++) D64[(l / 16 + 4 * r) * 16 + l % 16] = e[r];
}

void launch_mfma_probe(const float *A32, const float *B32, float *D32, const double *A64,
                       const double *B64, double *D64, const float *Abf, const float *Bbf, float *Dbf,
                       hipStream_t s)
{
    mfma_probe_kernel<<<1, 64, 0, s>>>(A32, B32, D32, A64, B64, D64, Abf, Bbf, Dbf);
}

}  // namespace qjx
