// qjx_step.hip -- one propagation iteration over a table of ensembles (gfx950).
//
// Every active trajectory of every ensemble in the table advances by exactly one
// matrix-vector step at its own (branch, level) of the bisection tree
// (reference qj_proc.cpp:271-343):
//   plan   : prefixes over items (ensemble e, group g = b*D + lvl) of GEMM tiles and
//            list entries; clears the counters the iteration appends to.
//   gemm   : grouped GEMM  Y = U_g X  over each item's trajectories (columns gathered by
//            trajectory id), output to the spare slot, FP64 partial ||y||^2 per row
//            tile.  Persistent workgroups claim tiles from per-XCD ranges.
//   decide : FP64 crossing test and the recursion as a state machine; jumps queued.
//   jump   : recovery (experiment.cpp:1273-1382), one workgroup per jump.
// All arguments live in device memory (RunArgs points at them), so an iteration's
// launches are identical from one iteration to the next except for the parity
// and can be captured in a HIP graph.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "qjx_device.h"
#include "qjx_internal.h"
#include "qjx_kernels.h"

namespace qjx {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));
typedef double double2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

// Loads and stores of device buffers through the global address space: the
// pointers come out of the ensemble table, where the compiler cannot see their
// address space and would emit flat instructions, whose completion an LDS wait
// (lgkmcnt) also has to cover -- that would serialise every prefetch of the next K
// block with the fragment reads of the current one.
#define QJX_GLOBAL __attribute__((address_space(1)))
template <typename V, typename T>
__device__ inline V gld(const T *p)
{
    return *(const QJX_GLOBAL V *)(p);
}
template <typename V, typename T>
__device__ inline void gst(T *p, V v)
{
    *(QJX_GLOBAL V *)(p) = v;
}

// largest item with pre[item] <= t (pre non-decreasing, pre[n] > t)
// The FP64 GEMM is compiled in a translation unit of its own (qjx_step_f64.hip
// includes this file with QJX_STEP_F64 defined): the rest of this file is built with
// the scheduler that keeps memory clauses together, which speeds the FP32 and
// 3xBF16 GEMMs and slows the FP64 one.
void launch_gemm_f64(const RunArgs &a, const StepGrid &g, hipStream_t s);
hipError_t gemm_f64_occupancy(int *n, int bn);

__device__ inline int plan_search(const int *pre, int n, int t)
{
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// plan: one workgroup of 1024 threads, 4 consecutive items per thread (all of
// one ensemble, since kMaxGroups is a multiple of 4).
// ---------------------------------------------------------------------------
constexpr int kPlanThreads = 1024;
static_assert(kPlanItems == 4 * kPlanThreads, "plan layout");

#if !defined(QJX_STEP_F64)
__global__ __launch_bounds__(kPlanThreads) void plan_kernel(RunArgs a)
{
    __shared__ int s_t[kPlanThreads];
    __shared__ int s_e[kPlanThreads];
    const int tid = threadIdx.x;
    const int p = a.parity;
    const int item0 = 4 * tid;
    const int e = item0 / kMaxGroups;
    static_assert(kPlanItems / kMaxGroups <= kMaxEns, "plan ensembles");
    // the thread's ensemble record is loaded with the table size (the table holds
    // kMaxEns records; those past the size are not used)
    const Ens &E = a.ens[e];
    const int G = E.G, R = E.R;
    int *const cnt_cur = E.cnt[p], *const cnt_next = E.cnt[p ^ 1];
    const int nens = *a.nens;
    const int nitems = nens * kMaxGroups;
    int tl[4], el[4], tsum = 0, esum = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        tl[q] = el[q] = 0;
        if (item0 + q < nitems) {
            const int g = (item0 + q) % kMaxGroups;
            const int c = g < G ? cnt_cur[g] : 0;
            el[q] = c;
            tl[q] = ((c + a.bn - 1) / a.bn) * ((R + a.rc - 1) / a.rc);   // work units: column tiles x row chunks
            cnt_next[g] = 0;          // the decide kernel of this iteration appends here
        }
        tsum += tl[q];
        esum += el[q];
    }
    if (esum > 0) atomicAdd(&E.counters[0], (unsigned long long)esum);
    s_t[tid] = tsum;
    s_e[tid] = esum;
    __syncthreads();
    // inclusive Hillis-Steele scan over the threads that hold items
    int nthr = 1;
    while (nthr < (nitems + 3) / 4) nthr <<= 1;
    for (int off = 1; off < nthr; off <<= 1) {
        const int vt = tid >= off ? s_t[tid - off] : 0;
        const int ve = tid >= off ? s_e[tid - off] : 0;
        __syncthreads();
        s_t[tid] += vt;
        s_e[tid] += ve;
        __syncthreads();
    }
    int tb = s_t[tid] - tsum, eb = s_e[tid] - esum;   // exclusive bases
    Plan *pl = a.plan;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if (item0 + q < nitems) {
            pl->tile_pre[item0 + q] = tb;
            pl->ent_pre[item0 + q] = eb;
        }
        tb += tl[q];
        eb += el[q];
    }
    if (tid == 0) {
        pl->tile_pre[nitems] = s_t[nthr - 1];
        pl->ent_pre[nitems] = s_e[nthr - 1];
        pl->cols += (unsigned long long)s_e[nthr - 1];
        pl->next_total[p ^ 1] = 0;
        *a.jcnt = 0;
    }
    if (tid < 8) pl->xnext[tid] = 0;
}
#endif  // !QJX_STEP_F64

// Claim this persistent workgroup's next tile.  The tile space is cut into 8
// contiguous ranges, one per XCD (workgroup w runs on XCD w % 8), so the row tiles
// of a column tile are claimed on one XCD and share its L2; a workgroup whose range
// is exhausted takes tiles from the other XCDs' ranges.  Lane 0 of wave 0 claims
// with an atomic (tile_claim, issued early: the GEMM kernels claim their next tile
// before their epilogue's stores and wait for the answer after them); wave 0 then
// finds the tile's item in the plan prefix with a two-level ballot search, places
// the tile in its item and leaves the result in LDS for tile_read.  The search's
// first level is loaded once per launch (search_l1: the prefix does not change
// during one), and the second level comes with the item's column count (the entry
// prefix) and its ensemble's row-tile count: one dependent load round per claim.
struct TileRef {
    int e, g, rt, col0, ncol;
};

static_assert(kPlanItems <= 64 * 64, "two-level tile search");

// Lane 0: the next index of this workgroup's XCD range.  A claim past the range's
// end is harmless (the counters restart every iteration): tile_resolve then takes a
// tile from another range.
__device__ inline int tile_claim(const RunArgs &a)
{
    return threadIdx.x == 0 ? atomicAdd(&a.plan->xnext[blockIdx.x & 7], 1) : 0;
}

// Wave 0, once per launch: lane l's value of the search's first level, pre[64 l]
// (INT_MAX past the items).
__device__ inline int search_l1(const int *pre, int nitems)
{
    const int i1 = (threadIdx.x & 63) * 64;
    return i1 < nitems ? gld<int>(pre + i1) : 0x7fffffff;
}

// Wave 0: the tile of claim c0 (lane 0's tile_claim result) -> s_claim[0..4] =
// ensemble, group, row tile (row chunk of a.rc row tiles), first column, columns;
// s_claim[0] = -1 when no tile is left.  v1: this lane's search_l1.
template <int BN>
__device__ inline void tile_resolve(const RunArgs &a, const int *pre, int nitems, int total, int c0, int v1,
                                    int *s_claim)
{
    const int lane = threadIdx.x;
    int t = -1;
    if (lane == 0) {
        const int x0 = blockIdx.x & 7;
        for (int i = 0; i < 8 && t < 0; i++) {
            const int x = (x0 + i) & 7;
            const int lo = (int)(((long long)total * x) >> 3), hi = (int)(((long long)total * (x + 1)) >> 3);
            if (i == 0) {
                if (c0 < hi - lo) t = lo + c0;
                continue;
            }
            if (hi <= lo) continue;
            if (a.plan->xnext[x] >= hi - lo) continue;           // cheap pre-check
            const int c = atomicAdd(&a.plan->xnext[x], 1);
            if (c < hi - lo) t = lo + c;
        }
    }
    t = __shfl(t, 0, 64);
    if (t < 0) {
        if (lane == 0) s_claim[0] = -1;
        return;
    }
    // the largest item with pre[item] <= t (pre non-decreasing, pre[0] = 0,
    // pre[nitems] = total > t).  The 64 items from b are one ensemble's groups
    // (kMaxGroups = 64), so its row-tile count loads with the second level.
    static_assert(kMaxGroups == 64, "one ensemble per second-level range");
    const unsigned long long m1 = __ballot(v1 <= t);
    const int b = (63 - __clzll(m1)) * 64;
    const int i2 = b + lane;
    const int *ent = a.plan->ent_pre;
    const int R = a.ens[b / kMaxGroups].R;
    int v2 = 0x7fffffff, n0 = 0, n1 = 0;
    if (i2 < nitems) {
        v2 = gld<int>(pre + i2);
        n0 = gld<int>(ent + i2);
        n1 = gld<int>(ent + i2 + 1);
    }
    const unsigned long long m2 = __ballot(v2 <= t);
    const int w = 63 - __clzll(m2);
    const int item = b + w;
    const int local = t - __shfl(v2, w, 64);
    const int e = item / kMaxGroups, g = item % kMaxGroups;
    // tiles of an item: super-columns of a.sc column tiles, each visited row tile
    // by row tile (row chunk by row chunk: a.rc row tiles per unit), so the row
    // tile's propagator band is reused from L2 by the super-column's column tiles
    // and the column tiles' states by the row tiles.  The item's column count is
    // its list's entry count (entry prefix).
    const int cnt = __shfl(n1 - n0, w, 64);
    const int nct = (cnt + BN - 1) / BN;
    const int per = ((R + a.rc - 1) / a.rc) * a.sc;
    const int sc = local / per, rem = local - sc * per;
    const int cw = min(a.sc, nct - sc * a.sc);
    const int col0 = (sc * a.sc + rem % cw) * BN;
    if (lane == 0) {
        s_claim[0] = e;
        s_claim[1] = g;
        s_claim[2] = rem / cw;
        s_claim[3] = col0;
        s_claim[4] = min(BN, cnt - col0);
    }
}

// Every thread, after a barrier that follows tile_resolve: the claimed tile (held
// in scalar registers), false when none is left.
__device__ inline bool tile_read(const int *s_claim, TileRef &r)
{
    r.e = __builtin_amdgcn_readfirstlane(s_claim[0]);
    if (r.e < 0) return false;
    r.g = __builtin_amdgcn_readfirstlane(s_claim[1]);
    r.rt = __builtin_amdgcn_readfirstlane(s_claim[2]);
    r.col0 = __builtin_amdgcn_readfirstlane(s_claim[3]);
    r.ncol = __builtin_amdgcn_readfirstlane(s_claim[4]);
    return true;
}

// Claim and read in one step (kernels that do not overlap the claim with work).
template <int BN>
__device__ inline bool tile_next(const RunArgs &a, const int *pre, int nitems, int total, int *s_claim,
                                 TileRef &r)
{
    if (threadIdx.x < 64) tile_resolve<BN>(a, pre, nitems, total, tile_claim(a), search_l1(pre, nitems), s_claim);
    __syncthreads();
    const bool ok = tile_read(s_claim, r);
    __syncthreads();
    return ok;
}

// The two state slots of a tile's ensemble (wave-uniform: read once per tile)
// and the column of trajectory tr in slot cur.
struct Slots {
    char *p0, *p1;
    template <typename T>
    __device__ T *col(int cur, int tr, int L) const
    {
        return reinterpret_cast<T *>(cur ? p1 : p0) + (size_t)tr * L;
    }
};

__device__ inline Slots tile_slots(const Ens &E)
{
    return Slots{static_cast<char *>(E.slot[0]), static_cast<char *>(E.slot[1])};
}

template <int BN>
__device__ inline void load_cols(const RunArgs &a, const TileRef &r, int *s_tr, int *s_cur)
{
    const Ens &E = a.ens[r.e];
    const int tid = threadIdx.x;
    if (tid < BN) {
        // columns past ncol repeat the first one: loaded, never stored
        const int c = tid < r.ncol ? tid : 0;
        const int ent = gld<int>(E.list[a.parity] + (size_t)r.g * E.M + r.col0 + c);
        s_tr[tid] = ent >> 1;          // list entries are (trajectory << 1) | current slot
        s_cur[tid] = ent & 1;
    }
}

// ---------------------------------------------------------------------------
// Matrix-core operand layouts (checked by the probe kernel below):
//   v_mfma_f32_32x32x2f32: A (32 x 2): lane l holds A[l % 32][l / 32]; B (2 x 32):
//     B[l / 32][l % 32]; D (32 x 32): lane l, register r holds
//     D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32]
//   v_mfma_f32_32x32x16_bf16: lane l holds 8 consecutive k of row / column l % 32
//     (lane half h = l / 32 selects which 8); D as above
// Block-real tiles (FP64 form): 128 rows of the 2dp x 2dp operand, K blocks of 32.
// ---------------------------------------------------------------------------
constexpr int MF_BM = 128, MF_BK = 32;

// The K blocks a tile multiplies: [kb0, kb0 + nb) of the operand's first half,
// then the same blocks of its second half (offset dp).  The blocks outside hold
// exact zeros of the propagator's row tile (band table), so skipping them leaves
// every accumulator bit-identical; nb = 0 leaves the tile zero.
struct KRange {
    int kb0, nb, hb;
    __device__ int count() const { return 2 * nb; }
    __device__ int at(int i) const { return (i < nb ? kb0 + i : hb + kb0 + i - nb) * MF_BK; }
    __device__ int half_block(int i) const { return kb0 + (i < nb ? i : i - nb); }   // block within its half
};

__device__ inline KRange tile_krange(const Ens &E, const TileRef &r)
{
    const int *p = E.krange + 2 * ((size_t)r.g * E.R + r.rt);
    KRange k;
    k.kb0 = gld<int>(p);
    k.nb = gld<int>(p + 1) - k.kb0;
    k.hb = E.dp / MF_BK;
    return k;
}

// The K blocks each 32-row fragment of the tile needs (band table kfrag): a
// fragment's matrix products outside its range multiply exact zeros of the
// propagator and are skipped, which leaves its accumulators bit-identical.
struct FragRange {
    unsigned lo, hi;   // byte f: fragment f's first and past-the-end block within a half
    __device__ bool active(int f, int b) const
    {
        return b >= (int)((lo >> (8 * f)) & 0xffu) && b < (int)((hi >> (8 * f)) & 0xffu);
    }
    // fragment-blocks the tile multiplies over both halves of the operand
    __device__ unsigned work() const
    {
        unsigned w = 0;
#pragma unroll
        for (int f = 0; f < 4; f++) w += ((hi >> (8 * f)) & 0xffu) - ((lo >> (8 * f)) & 0xffu);
        return 2 * w;
    }
};

__device__ inline FragRange tile_frags(const Ens &E, const TileRef &r)
{
    const unsigned *p = E.kfrag + kFragWords * ((size_t)r.g * E.R + r.rt);
    FragRange f;
    f.lo = __builtin_amdgcn_readfirstlane(gld<unsigned>(p));
    f.hi = __builtin_amdgcn_readfirstlane(gld<unsigned>(p + 1));
    return f;
}

#if !defined(QJX_STEP_F64)
// 3 x BF16 LDS rows: 32 bf16 (64 bytes) with the 16-byte chunk index XOR-swizzled by
// (row / 4) % 4, which keeps the fragment reads conflict-free.
constexpr int BF_ROW = MF_BK;          // bf16 per LDS row, no padding

__device__ inline int bf_off(int row, int k)   // element offset of (row, k), k < 32
{
    const int chunk = (k >> 3) ^ ((row >> 2) & 3);
    return row * BF_ROW + chunk * 8 + (k & 7);
}


// ---------------------------------------------------------------------------
// Complex-pair GEMM (FP32 and 3xBF16 forms).  These forms store a propagator as
// dp rows [Re U_i,0 .. Re U_i,dp-1 | Im U_i,0 .. Im U_i,dp-1] (row-major dp x 2dp).
// The block-real operand holds every entry of U twice and one of its tiles yields
// either Re y or Im y; a complex-pair row tile is kCpRows = 64 complex rows and
// writes both Re y and Im y of its rows, so each entry of U is staged once per
// column tile and its K range is that of 64 rows instead of 128.  A K block is
// kCpBlock = 16 complex columns; an LDS row holds [Re 0..15 | Im 0..15] of one
// propagator row or one state column.  Wave (wm, wn) of the 2 x NWN waves takes the
// row tile's complex rows 32 wm .. 32 wm + 31 and BN / NWN columns.  The FP64 partial
// norm of a row tile covers Re and Im of its 64 rows; R = dp / 64 equals the
// block-real tile count 2 dp / 128, so the decide kernel is shared.
//
// Work unit: one column tile and up to a.rc row tiles (strided: unit c takes row
// tiles c, c + nch, ...), walked as one stream of K blocks.  The register stages and
// LDS buffers run on across the row tiles' boundaries, so the next row tile's first
// blocks are already loaded while the current one finishes, and the column gather,
// the claim and the first loads are paid once per unit instead of once per row
// tile.  At a row tile's last block its outputs and norm partials are written and
// the accumulators restart.
//
// FP32 (FM = kCpF32x16): Gauss's three-product complex multiply on
// v_mfma_f32_16x16x4f32.  With U = A + iB and x = a + ib:
//   k1 = (A + B) a,   k2 = A (b - a),   k3 = B (a + b),
//   Re y = k1 - k3,   Im y = k1 + k2,
// three real products per complex product instead of four; A + B, b - a and a + b
// are formed in registers from the fragments read out of LDS.  A wave's rows are
// two 16-row fragments, and each fragment multiplies only the 4-column k-steps
// inside its own band window: a 16-row fragment's window is narrower than a 32-row
// one's, and at 4 instead of 16 columns its rounding wastes less, so the form
// executes about a quarter fewer products than on 32-row fragments with 16-column
// blocks.  Within each 16-column part of an LDS row, column k sits at position
// 4 (k % 4) + k / 4, so lane group g of the 16x16x4 product reads k = g, g + 4,
// g + 8, g + 12 (one per k-step) with one 16-byte read, and k-step s covers the
// contiguous columns 4 s .. 4 s + 3.
//
// 3xBF16 (FM = kCpBF16x3): split products, x = hi + lo (hi = bf16(x), lo =
// bf16(x - hi)) for both operands and U X ~= Ulo Xhi + Uhi Xlo + Uhi Xhi (the
// propagator is split when packed, the state tile when staged into LDS), FP32
// accumulation, with four real products per complex product,
// Re y = A a - B b and Im y = B a + A b; -B is B with its sign bits flipped in
// registers.  Each accumulator takes its four small cross terms before its two
// hi x hi terms.  v_mfma_f32_32x32x16_bf16 takes one part (Re or Im: 16 k) of a K
// block, lane half h holding k = 8 h .. 8 h + 7; LDS rows are 64 bytes
// (Re 16 | Im 16 bf16), XOR-swizzled (bf_off).  A wave's 32 rows are one band
// fragment, skipped in the K blocks outside its range.
//
// The k order, and the row split over waves, do not depend on BN or on the unit,
// so results do not depend on where a column lands, on the tile width or on rc.
// ---------------------------------------------------------------------------
constexpr int kCpChunkMax = 8;               // row tiles per work unit at most (bytes of a 64-bit word)

typedef unsigned uintx4 __attribute__((ext_vector_type(4)));

// -x of a bf16 vector: the sign bits flipped (exact)
__device__ inline bf16x8 bf16_neg(bf16x8 v)
{
    uintx4 u = __builtin_bit_cast(uintx4, v);
    u ^= 0x80008000u;
    return __builtin_bit_cast(bf16x8, u);
}

__device__ inline int byte_at(uint64_t w, int i) { return (int)((w >> (8 * i)) & 0xffu); }

// A wave-uniform pointer read into scalar registers now (the builtin returns a signed
// int, so each half goes through uint32_t before it is widened)
template <typename T>
__device__ inline T *uniform_ptr(T *p)
{
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    return reinterpret_cast<T *>(((uint64_t)hi << 32) | lo);
}

template <int FM, int BN, int NWN>
__global__ __launch_bounds__(128 * NWN) __attribute__((amdgpu_waves_per_eu(2))) void gemm_cp(RunArgs a)
{
    static_assert(FM == kCpBF16x3 || FM == kCpF32x16, "form");
    constexpr bool BF = FM == kCpBF16x3;     // 3xBF16: 32x32x16 products, 32-row fragments
    constexpr bool F16 = FM == kCpF32x16;    // FP32: 16x16x4 products, 16-row fragments
    // FP32 with 8 waves (BN = 256): three LDS buffers and each block's fragments read
    // ahead in two 8-byte halves (RA), 34-float LDS rows, where 8-byte reads of 16
    // consecutive rows fall in distinct banks.  The 4-wave FP32 kernels keep two
    // buffers (their LDS then fits two workgroups per CU) and read a block's fragments
    // as it starts, 16 bytes at a time, from 36-float rows.
    constexpr bool RA = F16 && NWN == 4;
    constexpr int CP_PITCH = RA ? 2 * kCpBlock + 2 : 2 * kCpBlock + 4;
    constexpr int T = 128 * NWN;             // 2 x NWN waves
    constexpr int NJ = BN / (32 * NWN);      // 32-column MFMA tiles per wave (16-column: 2 NJ)
    constexpr int NJ16 = 2 * NJ;
    constexpr int RP = T / 8;                // LDS rows filled per pass (8 threads per row)
    constexpr int NQ = BN / RP;              // state columns copied per thread
    // propagator loads of a K block: 512 16-byte chunks (FP32: 64 rows x 8; BF16: 64
    // rows x 4 of the hi and of the lo array), NA per thread
    constexpr int NA = 2 * kCpRows * 4 / T;
    constexpr int NACC = BF ? 2 : 3;         // accumulator sets per column tile
    static_assert(NJ >= 1 && NA >= 1 && NQ >= 1 && NA * T == 2 * kCpRows * 4, "tile shape");
    using AV = std::conditional_t<BF, bf16x8, floatx4>;
    // LDS: two buffers of the propagator tile and the state tile
    constexpr int A_ELEMS = BF ? 2 * kCpRows * BF_ROW : kCpRows * CP_PITCH;   // per buffer (hi, lo)
    constexpr int B_ELEMS = BF ? 2 * BN * BF_ROW : BN * CP_PITCH;
    using LT = std::conditional_t<BF, __bf16, float>;
    // LDS buffers: three with read-ahead (a block is stored two iterations before its
    // products, so the next block's fragments can be read during this one's)
    constexpr int NBUF = RA ? 3 : 2;
    __shared__ __attribute__((aligned(16))) LT s_a[NBUF][A_ELEMS];
    __shared__ __attribute__((aligned(16))) LT s_b[NBUF][B_ELEMS];
    __shared__ int s_tr[BN];
    __shared__ int s_cur[BN];
    __shared__ double s_norm[2][2][BN];       // [row tile parity][wm][column]
    __shared__ unsigned s_fw[kCpChunkMax][4];  // 16-row fragment windows of the unit's row tiles
    const int *s_tpre = a.plan->tile_pre;     // read through L2: one search per unit
    const int nitems = *a.nens * kMaxGroups;
    const int total = s_tpre[nitems];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = wid / NWN, wn = wid % NWN;
    const int h = lane >> 5, li = lane & 31;  // 32x32 products: lane half, row / column
    const int q4 = lane >> 4, l16 = lane & 15;  // 16x16 products: lane group, row / column
    const int cc = tid & 7, cr = tid >> 3;    // 16-byte chunk of a state LDS row; row
    __shared__ int s_claim[5];
    TileRef r;
    const int v1 = wid == 0 ? search_l1(s_tpre, nitems) : 0;
    if (wid == 0) tile_resolve<BN>(a, s_tpre, nitems, total, tile_claim(a), v1, s_claim);
    __syncthreads();
    while (tile_read(s_claim, r)) {
        const Ens &E = a.ens[r.e];
        const Slots sl = tile_slots(E);
        const int L = E.L, dp = E.dp, R = E.R;
        // unit c of the column tile takes row tiles c, c + nch, c + 2 nch, ... (nch
        // units per column tile): the units claimed together then cover neighbouring
        // row tiles, which share state blocks (the band overlap) through L2, and a
        // unit's own row tiles share none (3xBF16 1.9% and FP32 0.5% faster than
        // units of adjacent row tiles)
        const int nch = (R + a.rc - 1) / a.rc;
        const int rt0 = r.rt;
        const int nrt = (R - 1 - rt0) / nch + 1;     // row tiles in the unit
        // the norm-partial array, read here into scalar registers: loaded lazily at a
        // row tile's end, its load would wait (in-order vmcnt) for every K block
        // prefetched ahead of it
        double *const normpart = uniform_ptr(E.normpart);
        // band of each row tile of the unit, one byte per row tile (wave-uniform):
        // first K block, block count, and (3xBF16) the two 32-row fragments' block
        // ranges; the FP32 form's 16-row fragment windows go to LDS
        uint64_t bk0 = 0, bnb = 0, bf0 = 0, bf1 = 0, bh0 = 0, bh1 = 0;
        unsigned work = 0;
#pragma unroll
        for (int i = 0; i < kCpChunkMax; i++) {
            if (i < nrt) {
                const size_t t = (size_t)r.g * R + rt0 + i * nch;
                const int k0 = __builtin_amdgcn_readfirstlane(gld<int>(E.krange + 2 * t));
                const int k1 = __builtin_amdgcn_readfirstlane(gld<int>(E.krange + 2 * t + 1));
                bk0 |= (uint64_t)k0 << (8 * i);
                bnb |= (uint64_t)(k1 - k0) << (8 * i);
                if (F16) {
#pragma unroll
                    for (int f = 0; f < 4; f++) {
                        const unsigned w = __builtin_amdgcn_readfirstlane(gld<unsigned>(E.kfrag + kFragWords * t + f));
                        work += (w >> 16) - (w & 0xffffu);
                        if (tid == 0) s_fw[i][f] = w;
                    }
                } else {
                    const unsigned lo = __builtin_amdgcn_readfirstlane(gld<unsigned>(E.kfrag + kFragWords * t));
                    const unsigned hi = __builtin_amdgcn_readfirstlane(gld<unsigned>(E.kfrag + kFragWords * t + 1));
                    bf0 |= (uint64_t)(lo & 0xffu) << (8 * i);
                    bf1 |= (uint64_t)((lo >> 8) & 0xffu) << (8 * i);
                    bh0 |= (uint64_t)(hi & 0xffu) << (8 * i);
                    bh1 |= (uint64_t)((hi >> 8) & 0xffu) << (8 * i);
                    work += ((hi & 0xffu) - (lo & 0xffu)) + (((hi >> 8) & 0xffu) - ((lo >> 8) & 0xffu));
                }
            }
        }
        // the unit's columns, gathered once the band table's loads are issued, so that
        // the two wait together
        load_cols<BN>(a, r, s_tr, s_cur);
        __syncthreads();
        int nbt = 0;                                 // K blocks of the unit
#pragma unroll
        for (int i = 0; i < kCpChunkMax; i++)
            if (i < nrt) nbt += byte_at(bnb, i);
        // this thread's load addresses at the unit's first row tile, block offset 0
        const int koff = cc < 4 ? 4 * cc : dp + 4 * (cc - 4);   // Re part (chunks 0..3) or Im part
        // LDS position of this thread's chunk (FP32): contiguous, or for 16-row
        // fragments the chunk's element e (k = 4 (cc % 4) + e) at 4 e + cc % 4 of its
        // part, so that lane group g of a 16x16x4 product reads k = g, g + 4, g + 8,
        // g + 12 (one per product step) with one 16-byte read
        const int lpos = F16 ? (cc >> 2) * kCpBlock + (cc & 3) : 4 * cc;
        const LT *aptr[NA];
        int aoff[NA];
        {
            const size_t PL = (size_t)dp * L;
            const LT *U = reinterpret_cast<const LT *>(E.props) + (size_t)r.g * PL;
#pragma unroll
            for (int q = 0; q < NA; q++) {
                if (BF) {
                    // load index e = tid + T q: array e / 256 (hi, lo; wave-uniform), row
                    // (e % 256) / 4, 8-element chunk e % 4 (0, 1: Re part; 2, 3: Im part)
                    const LT *Ul = reinterpret_cast<const LT *>(E.props_lo) + (size_t)r.g * PL;
                    const int e = tid + T * q, row = (e & 255) >> 2, ch = e & 3;
                    const int gk = ch < 2 ? 8 * ch : dp + 8 * (ch - 2);
                    aptr[q] = ((e >> 8) ? Ul : U) + (size_t)(rt0 * kCpRows + row) * L + gk;
                    aoff[q] = (e >> 8) * kCpRows * BF_ROW + bf_off(row, 8 * ch);
                } else {
                    const int row = cr + RP * q;
                    aptr[q] = U + (size_t)(rt0 * kCpRows + row) * L + koff;
                    aoff[q] = row * CP_PITCH + lpos;
                }
            }
        }
        const float *bcol[NQ];
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const int c = cr + RP * q;
            bcol[q] = sl.col<const float>(s_cur[c], s_tr[c], L) + koff;
        }
        // accumulators: 32x32 products (floatx16 per 32-column tile) or 16x16 products
        // (floatx4 per 16-row fragment and 16-column tile), the same registers
        floatx16 acc[NACC][NJ];
        auto zero_acc = [&]() {
#pragma unroll
            for (int p = 0; p < NACC; p++)
#pragma unroll
                for (int j = 0; j < NJ; j++)
#pragma unroll
                    for (int e = 0; e < 16; e++) acc[p][j][e] = 0.f;
        };
        // the 16x16 accumulator of product p, fragment f (of the wave's two) and
        // 16-column tile j, as a view of the same registers
        auto acc16 = [&](int p, int f, int j) -> floatx4 {
            const floatx16 &v = acc[p][(f * NJ16 + j) >> 2];
            const int o = 4 * ((f * NJ16 + j) & 3);
            return floatx4{v[o], v[o + 1], v[o + 2], v[o + 3]};
        };
        auto acc16_set = [&](int p, int f, int j, floatx4 x) {
            floatx16 &v = acc[p][(f * NJ16 + j) >> 2];
            const int o = 4 * ((f * NJ16 + j) & 3);
            v[o] = x[0];
            v[o + 1] = x[1];
            v[o + 2] = x[2];
            v[o + 3] = x[3];
        };
        zero_acc();
        struct Stage {
            AV a[NA];
            floatx4 b[NQ];
        } st0, st1;
        // load cursor (wave-uniform): row tile lrt of the unit, block lkb of its band;
        // past the unit's last block it stays there (reloads are in bounds, never stored)
        int lrt = 0, lkb = 0;
        auto gload = [&](Stage &sg) {
            const int k0 = (byte_at(bk0, lrt) + lkb) * kCpBlock;
            const size_t ao = (size_t)lrt * nch * kCpRows * L + k0;
#pragma unroll
            for (int q = 0; q < NA; q++) sg.a[q] = gld<AV>(aptr[q] + ao);
#pragma unroll
            for (int q = 0; q < NQ; q++) sg.b[q] = gld<floatx4>(bcol[q] + k0);
            if (lkb + 1 < byte_at(bnb, lrt)) {
                lkb++;
            } else if (lrt + 1 < nrt) {
                lrt++;
                lkb = 0;
            }
        };
        // one FP32 chunk into LDS: 16 bytes, or (F16) its four elements 4 apart
        auto st_chunk = [&](float *dst, floatx4 v) {
            if (F16) {
#pragma unroll
                for (int e = 0; e < 4; e++) dst[4 * e] = v[e];
            } else {
                *reinterpret_cast<floatx4 *>(dst) = v;
            }
        };
        auto lstore = [&](const Stage &sg, int buf) {
            if constexpr (BF) {
#pragma unroll
                for (int q = 0; q < NA; q++) *reinterpret_cast<AV *>(&s_a[buf][aoff[q]]) = sg.a[q];
            } else {
#pragma unroll
                for (int q = 0; q < NA; q++) st_chunk(&s_a[buf][aoff[q]], sg.a[q]);
            }
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const int rr = cr + RP * q;
                if constexpr (BF) {
                    bf16x4 h4, l4;
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        const float x = sg.b[q][e];
                        const __bf16 hb = (__bf16)x;
                        h4[e] = hb;
                        l4[e] = (__bf16)(x - (float)hb);
                    }
                    __bf16 *bb = reinterpret_cast<__bf16 *>(s_b[buf]);
                    *reinterpret_cast<bf16x4 *>(&bb[bf_off(rr, 4 * cc)]) = h4;
                    *reinterpret_cast<bf16x4 *>(&bb[BN * BF_ROW + bf_off(rr, 4 * cc)]) = l4;
                } else {
                    st_chunk(reinterpret_cast<float *>(s_b[buf]) + rr * CP_PITCH + lpos, sg.b[q]);
                }
            }
        };
        // the 16-row fragments' windows of the current row tile (F16), in units of 4
        // complex columns: the wave's fragments are 2 wm and 2 wm + 1
        unsigned fw0 = 0, fw1 = 0;
        auto load_fw = [&](int i) {
            if (F16) {
                fw0 = __builtin_amdgcn_readfirstlane(s_fw[i][2 * wm]);
                fw1 = __builtin_amdgcn_readfirstlane(s_fw[i][2 * wm + 1]);
            }
        };
        // FP32 with read-ahead: a block's fragments in two halves, k-steps 0, 1 (lo) and
        // 2, 3 (hi): the wave's two 16-row fragments (Re and Im of U) and its 16-column
        // tiles (a and b of x).  A block's hi half is read as its products start and its
        // lo half during the previous block's second half, so every LDS read overlaps
        // matrix products.
        float2 fr_lo[2], fi_lo[2], fr_hi[2], fi_hi[2];
        float2 xa_lo[NJ16], xb_lo[NJ16], xa_hi[NJ16], xb_hi[NJ16];
        auto read_half = [&](int buf, int half, float2 (&fr)[2], float2 (&fi)[2], float2 (&xa)[NJ16],
                             float2 (&xb)[NJ16]) {
            const float *sa = reinterpret_cast<const float *>(s_a[buf]);
            const float *sb = reinterpret_cast<const float *>(s_b[buf]);
#pragma unroll
            for (int f = 0; f < 2; f++) {
                const float *pa = &sa[(wm * 32 + f * 16 + l16) * CP_PITCH + 4 * q4 + 2 * half];
                fr[f] = *reinterpret_cast<const float2 *>(pa);
                fi[f] = *reinterpret_cast<const float2 *>(pa + kCpBlock);
            }
#pragma unroll
            for (int j = 0; j < NJ16; j++) {
                const float *pb = &sb[(wn * (BN / NWN) + j * 16 + l16) * CP_PITCH + 4 * q4 + 2 * half];
                xa[j] = *reinterpret_cast<const float2 *>(pb);
                xb[j] = *reinterpret_cast<const float2 *>(pb + kCpBlock);
            }
        };
        // k-steps 2 half, 2 half + 1 of block b: step s of fragment f runs if 4 b + s
        // lies in the fragment's window [lo, hi) (units of 4 complex columns)
        auto f16_steps = [&](int b, int half, const float2 (&fr)[2], const float2 (&fi)[2],
                             const float2 (&xa)[NJ16], const float2 (&xb)[NJ16]) {
            const int s00 = min(max((int)(fw0 & 0xffffu) - 4 * b, 0), 4);
            const int s01 = min(max((int)(fw0 >> 16) - 4 * b, 0), 4);
            const int s10 = min(max((int)(fw1 & 0xffffu) - 4 * b, 0), 4);
            const int s11 = min(max((int)(fw1 >> 16) - 4 * b, 0), 4);
#pragma unroll
            for (int t = 0; t < 2; t++) {
                const int s = 2 * half + t;
#pragma unroll
                for (int f = 0; f < 2; f++) {
                    if (s < (f ? s10 : s00) || s >= (f ? s11 : s01)) continue;
                    const float as = fr[f][t] + fi[f][t];
#pragma unroll
                    for (int j = 0; j < NJ16; j++) {
                        const float bd = xb[j][t] - xa[j][t], bs = xa[j][t] + xb[j][t];
                        acc16_set(0, f, j, __builtin_amdgcn_mfma_f32_16x16x4f32(as, xa[j][t], acc16(0, f, j), 0, 0, 0));
                        acc16_set(1, f, j, __builtin_amdgcn_mfma_f32_16x16x4f32(fr[f][t], bd, acc16(1, f, j), 0, 0, 0));
                        acc16_set(2, f, j, __builtin_amdgcn_mfma_f32_16x16x4f32(fi[f][t], bs, acc16(2, f, j), 0, 0, 0));
                    }
                }
            }
        };
        // mid(): the next stage's LDS stores and loads, issued between the block's
        // matrix products so that they overlap the products instead of following them.
        // b: the block's index in its half of the operand (FP32: the k-step windows).
        // (3xBF16 and FP32 without read-ahead; with it the block runs in iter below.)
        auto mma_block = [&](int buf, int b, auto &&mid) {
            if constexpr (BF) {
                const __bf16 *sa = reinterpret_cast<const __bf16 *>(s_a[buf]);
                const __bf16 *sb = reinterpret_cast<const __bf16 *>(s_b[buf]);
                const int ra = wm * 32 + li;
                const int oR = bf_off(ra, 8 * h), oI = bf_off(ra, kCpBlock + 8 * h);
                const bf16x8 rh = *reinterpret_cast<const bf16x8 *>(&sa[oR]);
                const bf16x8 rl = *reinterpret_cast<const bf16x8 *>(&sa[kCpRows * BF_ROW + oR]);
                const bf16x8 ih = *reinterpret_cast<const bf16x8 *>(&sa[oI]);
                const bf16x8 il = *reinterpret_cast<const bf16x8 *>(&sa[kCpRows * BF_ROW + oI]);
                bf16x8 ah[NJ], al[NJ], bh[NJ], bl[NJ];
#pragma unroll
                for (int j = 0; j < NJ; j++) {
                    const int col = wn * (BN / NWN) + j * 32 + li;
                    const int oa = bf_off(col, 8 * h), ob = bf_off(col, kCpBlock + 8 * h);
                    ah[j] = *reinterpret_cast<const bf16x8 *>(&sb[oa]);
                    al[j] = *reinterpret_cast<const bf16x8 *>(&sb[BN * BF_ROW + oa]);
                    bh[j] = *reinterpret_cast<const bf16x8 *>(&sb[ob]);
                    bl[j] = *reinterpret_cast<const bf16x8 *>(&sb[BN * BF_ROW + ob]);
                }
                const bf16x8 nih = bf16_neg(ih), nil = bf16_neg(il);
                // the Re and Im accumulators of each column tile go round in turn, so
                // consecutive MFMAs never depend on each other
#pragma unroll
                for (int j = 0; j < NJ; j++) {
                    acc[0][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(rl, ah[j], acc[0][j], 0, 0, 0);
                    acc[1][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(il, ah[j], acc[1][j], 0, 0, 0);
                }
#pragma unroll
                for (int j = 0; j < NJ; j++) {
                    acc[0][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(rh, al[j], acc[0][j], 0, 0, 0);
                    acc[1][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ih, al[j], acc[1][j], 0, 0, 0);
                }
#pragma unroll
                for (int j = 0; j < NJ; j++) {
                    acc[0][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nil, bh[j], acc[0][j], 0, 0, 0);
                    acc[1][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(rl, bh[j], acc[1][j], 0, 0, 0);
                }
                mid();
#pragma unroll
                for (int j = 0; j < NJ; j++) {
                    acc[0][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nih, bl[j], acc[0][j], 0, 0, 0);
                    acc[1][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(rh, bl[j], acc[1][j], 0, 0, 0);
                }
#pragma unroll
                for (int j = 0; j < NJ; j++) {
                    acc[0][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(rh, ah[j], acc[0][j], 0, 0, 0);
                    acc[1][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ih, ah[j], acc[1][j], 0, 0, 0);
                }
#pragma unroll
                for (int j = 0; j < NJ; j++) {
                    acc[0][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nih, bh[j], acc[0][j], 0, 0, 0);
                    acc[1][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(rh, bh[j], acc[1][j], 0, 0, 0);
                }
            } else {
                // k-step s (0..3) of fragment f runs if 4 b + s lies in the fragment's
                // window [lo, hi) (units of 4 complex columns)
                const int s00 = min(max((int)(fw0 & 0xffffu) - 4 * b, 0), 4);
                const int s01 = min(max((int)(fw0 >> 16) - 4 * b, 0), 4);
                const int s10 = min(max((int)(fw1 & 0xffffu) - 4 * b, 0), 4);
                const int s11 = min(max((int)(fw1 >> 16) - 4 * b, 0), 4);
                const float *sa = reinterpret_cast<const float *>(s_a[buf]);
                const float *sb = reinterpret_cast<const float *>(s_b[buf]);
                floatx4 fre[2], fim[2], xa[NJ16], xb[NJ16];
#pragma unroll
                for (int f = 0; f < 2; f++) {
                    const float *pa = &sa[(wm * 32 + f * 16 + l16) * CP_PITCH + 4 * q4];
                    fre[f] = *reinterpret_cast<const floatx4 *>(pa);
                    fim[f] = *reinterpret_cast<const floatx4 *>(pa + kCpBlock);
                }
#pragma unroll
                for (int j = 0; j < NJ16; j++) {
                    const float *pb = &sb[(wn * (BN / NWN) + j * 16 + l16) * CP_PITCH + 4 * q4];
                    xa[j] = *reinterpret_cast<const floatx4 *>(pb);
                    xb[j] = *reinterpret_cast<const floatx4 *>(pb + kCpBlock);
                }
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    if (s == 2) mid();
#pragma unroll
                    for (int f = 0; f < 2; f++) {
                        if (s < (f ? s10 : s00) || s >= (f ? s11 : s01)) continue;
                        const float as = fre[f][s] + fim[f][s];
#pragma unroll
                        for (int j = 0; j < NJ16; j++) {
                            const float bd = xb[j][s] - xa[j][s], bs = xa[j][s] + xb[j][s];
                            acc16_set(0, f, j, __builtin_amdgcn_mfma_f32_16x16x4f32(as, xa[j][s], acc16(0, f, j), 0, 0, 0));
                            acc16_set(1, f, j, __builtin_amdgcn_mfma_f32_16x16x4f32(fre[f][s], bd, acc16(1, f, j), 0, 0, 0));
                            acc16_set(2, f, j, __builtin_amdgcn_mfma_f32_16x16x4f32(fim[f][s], bs, acc16(2, f, j), 0, 0, 0));
                        }
                    }
                }
            }
        };
        // outputs and norm partial of row tile i of the unit (after its last block)
        auto epilogue = [&](int i) {
            const int r0 = (rt0 + i * nch) * kCpRows;
            if constexpr (F16) {
                // 16x16 products: lane (g, c) of fragment f holds rows 4 g .. 4 g + 3 of the
                // fragment in column c of its 16-column tile
                double ns[NJ16];
#pragma unroll
                for (int j = 0; j < NJ16; j++) {
                    ns[j] = 0.0;
                    const int c = wn * (BN / NWN) + j * 16 + l16;
                    const bool ok = c < r.ncol;
                    float *y = sl.col<float>(s_cur[c] ^ 1, s_tr[c], L);
#pragma unroll
                    for (int f = 0; f < 2; f++) {
                        const int row = r0 + wm * 32 + f * 16 + 4 * q4;
                        const floatx4 k1 = acc16(0, f, j), k2 = acc16(1, f, j), k3 = acc16(2, f, j);
                        floatx4 re, im;
#pragma unroll
                        for (int e = 0; e < 4; e++) {
                            re[e] = k1[e] - k3[e];
                            im[e] = k1[e] + k2[e];
                        }
                        if (ok) {
                            gst<floatx4>(y + row, re);
                            gst<floatx4>(y + dp + row, im);
                        }
#pragma unroll
                        for (int e = 0; e < 4; e++) ns[j] += (double)re[e] * (double)re[e];
#pragma unroll
                        for (int e = 0; e < 4; e++) ns[j] += (double)im[e] * (double)im[e];
                    }
                }
#pragma unroll
                for (int j = 0; j < NJ16; j++) {
                    ns[j] += __shfl_xor(ns[j], 16, 64);
                    ns[j] += __shfl_xor(ns[j], 32, 64);
                }
                if (q4 == 0) {
#pragma unroll
                    for (int j = 0; j < NJ16; j++) s_norm[i & 1][wm][wn * (BN / NWN) + j * 16 + l16] = ns[j];
                }
            } else {
                double ns[NJ];
#pragma unroll
                for (int j = 0; j < NJ; j++) {
                    ns[j] = 0.0;
                    const int c = wn * (BN / NWN) + j * 32 + li;
                    const bool ok = c < r.ncol;
                    float *y = sl.col<float>(s_cur[c] ^ 1, s_tr[c], L);
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const int row = r0 + wm * 32 + 8 * q + 4 * h;
                        floatx4 re, im;
#pragma unroll
                        for (int e = 0; e < 4; e++) {
                            re[e] = acc[0][j][4 * q + e];
                            im[e] = acc[1][j][4 * q + e];
                        }
                        if (ok) {
                            gst<floatx4>(y + row, re);
                            gst<floatx4>(y + dp + row, im);
                        }
#pragma unroll
                        for (int e = 0; e < 4; e++) ns[j] += (double)re[e] * (double)re[e];
#pragma unroll
                        for (int e = 0; e < 4; e++) ns[j] += (double)im[e] * (double)im[e];
                    }
                }
#pragma unroll
                for (int j = 0; j < NJ; j++) ns[j] += __shfl_xor(ns[j], 32, 64);
                if (h == 0) {
#pragma unroll
                    for (int j = 0; j < NJ; j++) s_norm[i & 1][wm][wn * (BN / NWN) + j * 32 + li] = ns[j];
                }
            }
            zero_acc();
        };
        // the norm partial of row tile i, after a barrier that follows its epilogue
        auto write_norm = [&](int i) {
            if (tid < r.ncol)
                gst<double>(normpart + (size_t)s_tr[tid] * R + rt0 + i * nch, s_norm[i & 1][0][tid] + s_norm[i & 1][1][tid]);
        };
        // issued stage by stage (not interleaved), as the loop body expects them.
        // Two buffers: block 0 in LDS, blocks 1 and 2 in registers; read-ahead: blocks 0
        // and 1 in LDS (and block 0's lo fragments read), blocks 2 and 3 in registers.
        gload(st0);
        __builtin_amdgcn_sched_barrier(0);
        gload(st1);
        __builtin_amdgcn_sched_barrier(0);
        lstore(st0, 0);
        gload(st0);
        if constexpr (RA) {
            __builtin_amdgcn_sched_barrier(0);
            lstore(st1, 1);
            gload(st1);
        }
        __syncthreads();
        if constexpr (RA) read_half(0, 0, fr_lo, fi_lo, xa_lo, xb_lo);
        load_fw(0);
        // MMA cursor (wave-uniform): row tile mrt, block mkb of its band; pend = row
        // tile whose norm partial waits in s_norm
        int mrt = 0, mkb = 0, pend = -1, cn = 0;
        // iteration f: multiply block f and, half-way through its products, store block
        // f + 1 into the other LDS buffer (last read before the previous barrier) and
        // refill its registers with block f + 3.  The stores drain while the second
        // half of the products runs, and the next block's fragment reads find no stores
        // ahead of them (against storing before the products: 3xBF16 1.8% and FP32
        // 1.6% faster; against storing after them: FP32 1.6% faster).  Read-ahead: buffer
        // f % 3 holds block f; half-way, block f + 2 is stored into buffer (f + 2) % 3
        // (block f - 1's, whose reads ended before the previous barrier), its registers
        // take block f + 4, and block f + 1's lo fragments are read (stored one
        // iteration ago, visible since the previous barrier).
        auto iter = [&](Stage &nxt, int f) {
            const int buf = RA ? f % 3 : f & 1;
            // the next unit is claimed three blocks before this one's end (its atomic has
            // returned by then) and resolved after the last block
            if (f == (nbt > 3 ? nbt - 3 : 0) && wid == 0) cn = tile_claim(a);
            // keep the next stage's conversions out of this iteration's products:
            // hoisted, they would wait for loads that are meant to stay in flight
            __builtin_amdgcn_sched_barrier(0);
            const int b = byte_at(bk0, mrt) + mkb;
            auto mid = [&]() {
                __builtin_amdgcn_sched_barrier(0);
                if constexpr (RA) {
                    lstore(nxt, buf == 0 ? 2 : buf - 1);
                    gload(nxt);
                    read_half(buf == 2 ? 0 : buf + 1, 0, fr_lo, fi_lo, xa_lo, xb_lo);
                } else {
                    lstore(nxt, buf ^ 1);
                    gload(nxt);
                }
                __builtin_amdgcn_sched_barrier(0);
            };
            // the wave's fragments multiply only the blocks inside their band ranges
            bool act;
            if constexpr (F16) {
                act = (4 * b + 4 > (int)(fw0 & 0xffffu) && 4 * b < (int)(fw0 >> 16)) ||
                      (4 * b + 4 > (int)(fw1 & 0xffffu) && 4 * b < (int)(fw1 >> 16));
            } else {
                const int fl = wm ? byte_at(bf1, mrt) : byte_at(bf0, mrt);
                const int fh = wm ? byte_at(bh1, mrt) : byte_at(bh0, mrt);
                act = b >= fl && b < fh;
            }
            if constexpr (RA) {
                if (act) read_half(buf, 1, fr_hi, fi_hi, xa_hi, xb_hi);
                __builtin_amdgcn_sched_barrier(0);
                if (act) f16_steps(b, 0, fr_lo, fi_lo, xa_lo, xb_lo);
                mid();
                if (act) f16_steps(b, 1, fr_hi, fi_hi, xa_hi, xb_hi);
            } else {
                if (act) mma_block(buf, b, mid);
                else mid();
            }
            __syncthreads();
            if (pend >= 0) {
                write_norm(pend);
                pend = -1;
            }
            if (mkb + 1 == byte_at(bnb, mrt)) {
                epilogue(mrt);
                pend = mrt;
                mrt++;
                mkb = 0;
                if (mrt < nrt) load_fw(mrt);
            } else {
                mkb++;
            }
        };
        // the stage holding the block stored next: block f + 1, with read-ahead f + 2
        Stage &se = RA ? st0 : st1, &so = RA ? st1 : st0;
        int f = 0;
        for (; f + 1 < nbt; f += 2) {
            iter(se, f);
            iter(so, f + 1);
        }
        if (f < nbt) iter(se, f);
        if (wid == 0) tile_resolve<BN>(a, s_tpre, nitems, total, cn, v1, s_claim);
        __syncthreads();
        write_norm(pend);
        // executed matrix-core flops in units of 128: FP32 3 products x 2 x 16 x 4 per
        // (16-row fragment, 4-column k-step); 3xBF16 (FP32-equivalent) 4 products x 2 x
        // 32 x 16 per (32-row fragment, K block)
        if (tid == 0) atomicAdd(&E.counters[2], (unsigned long long)(BF ? 32 : 3) * r.ncol * work);
        __syncthreads();
    }
}

#endif  // !QJX_STEP_F64
#if defined(QJX_STEP_F64)
// ---------------------------------------------------------------------------
// GEMM tile, FP64: v_mfma_f64_16x16x4f64
//   A (16 x 4): lane l holds A[l % 16][l / 16]; B (4 x 16): B[l / 16][l % 16]
//   D (16 x 16): lane l, register r holds D[(l / 16) + 4 r][l % 16]
// 8 waves as 4 x 2, wave tile 32 x BN/2, WG tile 128 x BN (BN = 128, 64 or 32); within
// a K block of 32, MFMA step s (0..7) takes k = s + 8 q for lane quarter q.
// ---------------------------------------------------------------------------
constexpr int F64_PITCH = MF_BK + 2;

template <int BN>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(BN <= 64 ? 4 : 1))) void gemm_f64(RunArgs a)
{
    constexpr int NJ = BN / 32;          // 16-column MFMA tiles per wave
    constexpr int NQ = BN / 32;          // B-tile rows (columns of X) copied per thread
    __shared__ int s_tr[BN];
    __shared__ int s_cur[BN];
    __shared__ __attribute__((aligned(16))) double As[MF_BM * F64_PITCH];
    __shared__ __attribute__((aligned(16))) double Bs[BN * F64_PITCH];
    __shared__ double s_norm[4][BN];
    const int *s_tpre = a.plan->tile_pre;     // read through L2: one search per tile
    const int nitems = *a.nens * kMaxGroups;
    const int total = s_tpre[nitems];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    const int qq = lane >> 4, li = lane & 15;
    const int cp_k = (tid & 15) * 2, cp_r = tid >> 4;
    __shared__ int s_claim[5];
    TileRef r;
    const int v1 = wid == 0 ? search_l1(s_tpre, nitems) : 0;
    if (wid == 0) tile_resolve<BN>(a, s_tpre, nitems, total, tile_claim(a), v1, s_claim);
    __syncthreads();
    while (tile_read(s_claim, r)) {
        load_cols<BN>(a, r, s_tr, s_cur);
        __syncthreads();
        const Ens &E = a.ens[r.e];
        const Slots sl = tile_slots(E);
        const int L = E.L, r0 = r.rt * MF_BM;
        const double *U = reinterpret_cast<const double *>(E.props) + (size_t)r.g * L * L;
        const double *arow = U + (size_t)(r0 + cp_r) * L + cp_k;
        const size_t astep = (size_t)32 * L;
        const double *bcol[NQ];
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const int c = cp_r + 32 * q;
            bcol[q] = sl.col<const double>(s_cur[c], s_tr[c], L) + cp_k;
        }
        doublex4 acc[2][NJ];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < NJ; j++) acc[i][j] = doublex4{0.0, 0.0, 0.0, 0.0};
        double2v ra[4], rb[NQ];
        auto gload = [&](int k0) {
#pragma unroll
            for (int q = 0; q < 4; q++) ra[q] = gld<double2v>(arow + q * astep + k0);
#pragma unroll
            for (int q = 0; q < NQ; q++) rb[q] = gld<double2v>(bcol[q] + k0);
        };
        auto lstore = [&]() {
#pragma unroll
            for (int q = 0; q < 4; q++)
                *reinterpret_cast<double2v *>(&As[(cp_r + 32 * q) * F64_PITCH + cp_k]) = ra[q];
#pragma unroll
            for (int q = 0; q < NQ; q++)
                *reinterpret_cast<double2v *>(&Bs[(cp_r + 32 * q) * F64_PITCH + cp_k]) = rb[q];
        };
        // (single-buffered: at one workgroup per CU a second buffer measured slower)
        const double *pa0 = &As[(wm * 32 + li) * F64_PITCH + 8 * qq];
        const double *pb0 = &Bs[(wn * (BN / 2) + li) * F64_PITCH + 8 * qq];
        const KRange kr = tile_krange(E, r);
        const int nk = kr.count();
        const FragRange fr = tile_frags(E, r);      // the wave's 32 rows are fragment wm
        if (nk > 0) {
            gload(kr.at(0));
            lstore();
        }
        __syncthreads();
        for (int kb = 0; kb < nk; kb++) {
            const bool more = kb + 1 < nk;
            if (more) gload(kr.at(kb + 1));
            const bool act = fr.active(wm, kr.half_block(kb));
#pragma unroll
            for (int half = 0; half < 2; half++) {
                if (!act) break;
                double2v fa[2][2], fb[NJ][2];
#pragma unroll
                for (int i = 0; i < 2; i++) {
                    const double *pa = pa0 + i * 16 * F64_PITCH + 4 * half;
                    fa[i][0] = *reinterpret_cast<const double2v *>(pa);
                    fa[i][1] = *reinterpret_cast<const double2v *>(pa + 2);
                }
#pragma unroll
                for (int j = 0; j < NJ; j++) {
                    const double *pb = pb0 + j * 16 * F64_PITCH + 4 * half;
                    fb[j][0] = *reinterpret_cast<const double2v *>(pb);
                    fb[j][1] = *reinterpret_cast<const double2v *>(pb + 2);
                }
#pragma unroll
                for (int s4 = 0; s4 < 4; s4++)
#pragma unroll
                    for (int i = 0; i < 2; i++)
#pragma unroll
                        for (int j = 0; j < NJ; j++)
                            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i][s4 >> 1][s4 & 1],
                                                                              fb[j][s4 >> 1][s4 & 1], acc[i][j],
                                                                              0, 0, 0);
            }
            __syncthreads();
            if (more) {
                lstore();
                __syncthreads();
            }
        }
        // the next tile is claimed now and resolved after this tile's stores
        int cn = 0;
        if (wid == 0) cn = tile_claim(a);
        double ns[NJ];
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            ns[j] = 0.0;
            const int c = wn * (BN / 2) + j * 16 + li;
            const bool ok = c < r.ncol;
            double *y = sl.col<double>(s_cur[c] ^ 1, s_tr[c], L);
#pragma unroll
            for (int i = 0; i < 2; i++) {
                const int row = r0 + wm * 32 + i * 16 + qq;
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    if (ok) gst<double>(y + row + 4 * e, acc[i][j][e]);
                    ns[j] += acc[i][j][e] * acc[i][j][e];
                }
            }
        }
        if (wid == 0) tile_resolve<BN>(a, s_tpre, nitems, total, cn, v1, s_claim);
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            ns[j] += __shfl_xor(ns[j], 16, 64);
            ns[j] += __shfl_xor(ns[j], 32, 64);
        }
        if (qq == 0) {
#pragma unroll
            for (int j = 0; j < NJ; j++) s_norm[wm][wn * (BN / 2) + j * 16 + li] = ns[j];
        }
        __syncthreads();
        if (tid < r.ncol)
            gst<double>(E.normpart + (size_t)s_tr[tid] * E.R + r.rt,
                        ((s_norm[0][tid] + s_norm[1][tid]) + s_norm[2][tid]) + s_norm[3][tid]);
        if (tid == 0) atomicAdd(&E.counters[2], 16ull * r.ncol * fr.work());   // 2048 flops per unit of work()
        __syncthreads();
    }
}

void launch_gemm_f64(const RunArgs &a, const StepGrid &g, hipStream_t s)
{
    if (a.bn == 32) gemm_f64<32><<<g.gemm, 512, 0, s>>>(a);
    else if (a.bn == 64) gemm_f64<64><<<g.gemm, 512, 0, s>>>(a);
    else gemm_f64<128><<<g.gemm, 512, 0, s>>>(a);
}

hipError_t gemm_f64_occupancy(int *n, int bn)
{
    return bn == 32   ? hipOccupancyMaxActiveBlocksPerMultiprocessor(n, gemm_f64<32>, 512, 0)
           : bn == 64 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(n, gemm_f64<64>, 512, 0)
                      : hipOccupancyMaxActiveBlocksPerMultiprocessor(n, gemm_f64<128>, 512, 0);
}
#endif  // QJX_STEP_F64
#if !defined(QJX_STEP_F64)
// ---------------------------------------------------------------------------
// GEMM tile, vector FMA (reference form for checks and tiny sizes): 64 x 64 tile,
// 4 x 4 register tile per thread, K block 16.
// ---------------------------------------------------------------------------
constexpr int FM_BM = 64, FM_BN = 64, FM_BK = 16;

template <typename T>
__global__ __launch_bounds__(256) void gemm_fma(RunArgs a)
{
    __shared__ int s_tr[FM_BN];
    __shared__ int s_cur[FM_BN];
    __shared__ T As[FM_BK][FM_BM + 4];
    __shared__ T Bs[FM_BK][FM_BN + 1];
    __shared__ double s_norm[FM_BM / 4][FM_BN + 1];
    const int *s_tpre = a.plan->tile_pre;     // read through L2: one search per tile
    const int nitems = *a.nens * kMaxGroups;
    const int total = s_tpre[nitems];
    const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
    const int a_i = tid / (FM_BK / 4), a_k = (tid % (FM_BK / 4)) * 4;
    const int b_j = tid / (FM_BK / 4), b_k = (tid % (FM_BK / 4)) * 4;
    __shared__ int s_claim[5];
    TileRef r;
    while (tile_next<FM_BN>(a, s_tpre, nitems, total, s_claim, r)) {
        load_cols<FM_BN>(a, r, s_tr, s_cur);
        __syncthreads();
        const Ens &E = a.ens[r.e];
        const Slots sl = tile_slots(E);
        const int L = E.L, r0 = r.rt * FM_BM;
        const T *U = reinterpret_cast<const T *>(E.props) + (size_t)r.g * L * L;
        const T *xcol = sl.col<const T>(s_cur[b_j], s_tr[b_j], L);
        T acc[4][4];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = T(0);
        for (int k0 = 0; k0 < L; k0 += FM_BK) {
            const T *src = U + (size_t)(r0 + a_i) * L + k0 + a_k;
#pragma unroll
            for (int q = 0; q < 4; q++) As[a_k + q][a_i] = src[q];
#pragma unroll
            for (int q = 0; q < 4; q++) Bs[b_k + q][b_j] = xcol[k0 + b_k + q];
            __syncthreads();
#pragma unroll
            for (int kk = 0; kk < FM_BK; kk++) {
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
            if (c < r.ncol) {
                T *y = sl.col<T>(s_cur[c] ^ 1, s_tr[c], L) + r0 + ty * 4;
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
        if (tid < r.ncol) {
            double s = 0.0;
            for (int q = 0; q < FM_BM / 4; q++) s += s_norm[q][tid];
            E.normpart[(size_t)s_tr[tid] * E.R + r.rt] = s;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// decide: after a completed step at level l (accepted, or the finest-level step
// that carried a jump; t has already advanced by dt_l): counter++ and, while levels
// finish, return to the parent (qj_proc.cpp:325-342: t -= dt_parent, then counter++
// and t += dt_parent).  Level 0 finishes after S/2 steps and closes the half period;
// a level l >= 1 finishes after 2 steps.
// ---------------------------------------------------------------------------
template <typename DT>
__device__ inline void complete_step(TrajState &s, int l, DT dt, int S)
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
        s.t -= dt(l);
        s.t += dt(l);
    }
    s.lvl = l;
}

// Appends are aggregated per workgroup: LDS counters for the groups of a window of
// two ensembles (the workgroup's entries are sorted by item, so they almost always
// fall in it), one global atomic per non-empty counter; entries outside the window
// reserve their slot directly.
constexpr int kWin = 2 * kMaxGroups;

__global__ __launch_bounds__(256) void decide_kernel(RunArgs a)
{
    __shared__ int s_pre[kPlanItems + 1];
    __shared__ int s_cnt[kWin + 3];      // [kWin] jumps, [kWin+1..2] jumps per window ensemble
    __shared__ int s_base[kWin + 1];
    __shared__ int s_next;
    __shared__ double s_dt[2][kMaxGroups / 2];   // step sizes of the window's ensembles
    static_assert(kPlanItems >= 256, "first prefix block");
    const Plan *pl = a.plan;
    const int p = a.parity;
    // issued together, so that they wait for memory once: the table size, the first
    // 256 prefix entries (all of them up to three ensembles; the array holds
    // kPlanItems + 1) and ensemble 0's arrays, which most entries use
    const int nens = *a.nens;
    s_pre[threadIdx.x] = pl->ent_pre[threadIdx.x];
    const Ens &E0 = a.ens[0];
    const int *const list0 = E0.list[p];
    TrajState *const st0 = E0.st;
    const double *const np0 = E0.normpart;
    const int M0 = E0.M, R0 = E0.R;
    // what the crossing test reads of the ensemble, held from the same loads
    struct Fields {
        const double *dt;
        double norm_tol;
        uint64_t stride;
        int D, S, jump, is_obs, logcap;
    };
    const Fields F0 = {E0.dt, E0.norm_tol, E0.stride, E0.D, E0.S, E0.jump, E0.is_obs, E0.log.cap};
    const int nitems = nens * kMaxGroups;
    for (int i = threadIdx.x + blockDim.x; i <= nitems; i += blockDim.x) s_pre[i] = pl->ent_pre[i];
    for (int i = threadIdx.x; i < kWin + 3; i += blockDim.x) s_cnt[i] = 0;
    if (threadIdx.x == 0) s_next = 0;
    __syncthreads();
    const int total = s_pre[nitems];
    const int base0 = blockIdx.x * blockDim.x;
    if (base0 >= total) return;                       // whole workgroup idle
    const int idx = base0 + threadIdx.x;
    const int e_lo = plan_search(s_pre, nitems, base0) / kMaxGroups;
    if (threadIdx.x < kMaxGroups) {
        // read while the entries' states are in flight; used after the barrier below
        const int w = threadIdx.x / (kMaxGroups / 2), l = threadIdx.x % (kMaxGroups / 2);
        if (e_lo + w < nens) {
            const Ens &W = a.ens[e_lo + w];
            if (l < W.D) s_dt[w][l] = W.dt[l];
        }
    }
    int e = -1, tr = -1, key = -1, local = -1, direct = -1, jlocal = -1, jsrc = 0, ncur = 0;
    double n = 0.0;
    TrajState s;
    Fields F = F0;
    const int *list = list0;
    TrajState *st = st0;
    const double *npart = np0;
    int M = M0, R = R0;
    if (idx < total) {
        const int item = plan_search(s_pre, nitems, idx);
        e = item / kMaxGroups;
        const int g = item % kMaxGroups;
        if (e != 0) {
            const Ens &E = a.ens[e];
            list = E.list[p];
            st = E.st;
            npart = E.normpart;
            M = E.M;
            R = E.R;
            F = {E.dt, E.norm_tol, E.stride, E.D, E.S, E.jump, E.is_obs, E.log.cap};
        }
        tr = list[(size_t)g * M + (idx - s_pre[item])] >> 1;
        s = st[tr];
        // the R partials (contiguous) are loaded together, then added in ascending
        // row-tile order
        const double *np = npart + (size_t)tr * R;
        if ((R & 1) == 0) {
            // even R: the partials start 16-byte aligned and come in pairs
            for (int r0 = 0; r0 < R; r0 += 16) {
                double2 v[8];
#pragma unroll
                for (int q = 0; q < 8; q++)
                    v[q] = r0 + 2 * q < R ? *reinterpret_cast<const double2 *>(np + r0 + 2 * q) : make_double2(0.0, 0.0);
#pragma unroll
                for (int q = 0; q < 8; q++)
                    if (r0 + 2 * q < R) {
                        n += v[q].x;
                        n += v[q].y;
                    }
            }
        } else {
            for (int r0 = 0; r0 < R; r0 += 16) {
                double v[16];
#pragma unroll
                for (int q = 0; q < 16; q++) v[q] = r0 + q < R ? np[r0 + q] : 0.0;
#pragma unroll
                for (int q = 0; q < 16; q++)
                    if (r0 + q < R) n += v[q];
            }
        }
    }
    __syncthreads();
    if (idx < total) {
        const Ens &E = a.ens[e];
        const bool win = e - e_lo < 2;
        const double *wdt = s_dt[win ? e - e_lo : 0];
        const double *gdt = F.dt;
        auto dt_at = [&](int l) { return win ? wdt[l] : gdt[l]; };
        const int lvl = s.lvl;
        const double dt = dt_at(lvl);
        bool alive = true;
        if (n > 1.0 + F.norm_tol) {
            // experiment.cpp:1247-1251 throws: record the first offender, stop it.
            if (atomicCAS(&E.err[0], 0, 1) == 0) {
                E.err[1] = tr;
                const unsigned long long bits = (unsigned long long)__double_as_longlong(n);
                E.err[2] = (int)(bits & 0xffffffffULL);
                E.err[3] = (int)(bits >> 32);
            }
            alive = false;
        } else {
            const bool crossed = !(n > s.eta);              // experiment.cpp:1253-1260
            if (lvl == F.D - 1) {
                if (crossed) {
                    if (F.jump > 0 && F.is_obs == 1) {
                        if (s.logn < F.logcap) {
                            const size_t o = (size_t)tr * E.log.cap + s.logn;
                            E.log.time[o] = s.t + dt;
                            E.log.norm[o] = n;
                            E.log.eta[o] = s.eta;
                            E.log.type[o] = 0;
                        }
                        s.logn++;
                        s.njumps++;
                    }
                    // recovery draws `ran` (one channel, so channel 0 is chosen),
                    // then eta is redrawn while it is 0 (qj_proc.cpp:310-314).
                    (void)mcg_next(s.rng, F.stride);
                    double ev = mcg_next(s.rng, F.stride);
                    while (ev == 0.0) ev = mcg_next(s.rng, F.stride);
                    s.eta = ev;
                    // the jump kernel reads the step just taken (slot cur ^ 1) and
                    // writes the jumped state over the pre-step state (slot cur)
                    jsrc = s.cur ^ 1;
                    jlocal = atomicAdd(&s_cnt[kWin], 1);
                    if (e - e_lo < 2) atomicAdd(&s_cnt[kWin + 1 + (e - e_lo)], 1);
                    else atomicAdd(&E.counters[1], 1ULL);
                } else {
                    s.cur ^= 1;
                }
                s.t += dt;
                complete_step(s, lvl, dt_at, F.S);
            } else if (crossed) {
                s.lvl = lvl + 1;          // bisect from the pre-step state
            } else {
                s.cur ^= 1;
                s.t += dt;
                complete_step(s, lvl, dt_at, F.S);
            }
        }
        st[tr] = s;
        ncur = s.cur;            // the jump kernel writes into this slot, too
        if (alive && s.pos < s.target) {
            const int gn = s.b * F.D + s.lvl;
            key = (e - e_lo) * kMaxGroups + gn;
            atomicAdd(&s_next, 1);
            if (e - e_lo < 2) local = atomicAdd(&s_cnt[key], 1);
            else direct = atomicAdd(&E.cnt[p ^ 1][gn], 1);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < kWin; t += blockDim.x) {
        const int c = s_cnt[t];
        const int ee = e_lo + t / kMaxGroups;
        s_base[t] = (c && ee < nens) ? atomicAdd(&a.ens[ee].cnt[p ^ 1][t % kMaxGroups], c) : 0;
    }
    if (threadIdx.x == 0) {
        s_base[kWin] = s_cnt[kWin] ? atomicAdd(a.jcnt, s_cnt[kWin]) : 0;
        if (s_next) atomicAdd(&a.plan->next_total[p ^ 1], s_next);
        for (int w = 0; w < 2; w++)
            if (s_cnt[kWin + 1 + w]) atomicAdd(&a.ens[e_lo + w].counters[1], (unsigned long long)s_cnt[kWin + 1 + w]);
    }
    __syncthreads();
    if (key >= 0) {
        const Ens &E = a.ens[e];
        const int gn = key % kMaxGroups;
        const int pos = local >= 0 ? s_base[key] + local : direct;
        E.list[p ^ 1][(size_t)gn * E.M + pos] = (tr << 1) | ncur;
    }
    if (jlocal >= 0) {
        JumpEnt je;
        je.ens = e;
        je.tr = tr;
        je.src = jsrc;
        je.pad = 0;
        je.n = n;
        a.jl[s_base[kWin] + jlocal] = je;
    }
}

// ---------------------------------------------------------------------------
// jump: recovery (experiment.cpp:1273-1382) for the single dimer dissipator, one
// workgroup per queued jump (persistent):
//   phi <- phi / ||phi||;  v = L phi;  phi <- v / sqrt(||v||^2)      (FP64)
// L is tridiagonal: L_kk = Ldiag[k], L_{k+1,k} = -Loff[k], L_{k,k+1} = +Loff[k];
// v_k sums its three terms in ascending column order.  Thread t owns the k = t + 256 m
// (m < PER, PER = ceil(d / 256) rounded up to 4 or 8) and issues every global load
// of its elements -- state and operator entries -- before it uses the first, so a
// jump waits for memory once.
// ---------------------------------------------------------------------------
constexpr int kJumpThreads = 256;

// LDS holds the normalised state: 2 x (jdim + 2) doubles, jdim = the largest d
// of the table (dynamic, so small models keep more workgroups resident).
template <typename T, int PER>
__global__ __launch_bounds__(kJumpThreads) void jump_kernel(RunArgs a, int jdim)
{
    extern __shared__ double jump_lds[];
    double *pre = jump_lds, *pim = jump_lds + (jdim + 2);
    __shared__ double red[kJumpThreads / 64];
    // issued together, so that a workgroup's first jump waits for memory once more
    // before its state loads: the jump count, the workgroup's first entry (in bounds:
    // the grid is never larger than the list's capacity) and ensemble 0's record
    const int njump = *a.jcnt;
    JumpEnt je = a.jl[blockIdx.x];
    const Ens &E0 = a.ens[0];
    const int d0 = E0.d, dp0 = E0.dp, L0 = E0.L;
    const double *const Ld0 = E0.Ldiag, *const Lo0 = E0.Loff;
    char *const sl00 = static_cast<char *>(E0.slot[0]), *const sl01 = static_cast<char *>(E0.slot[1]);
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    for (int j = blockIdx.x; j < njump; j += gridDim.x) {
        if (j != (int)blockIdx.x) je = a.jl[j];
        int d = d0, dp = dp0, L = L0;
        const double *Ldiag = Ld0, *Loff = Lo0;
        char *s0 = sl00, *s1 = sl01;
        if (je.ens != 0) {
            const Ens &E = a.ens[je.ens];
            d = E.d;
            dp = E.dp;
            L = E.L;
            Ldiag = E.Ldiag;
            Loff = E.Loff;
            s0 = static_cast<char *>(E.slot[0]);
            s1 = static_cast<char *>(E.slot[1]);
        }
        const double sn = sqrt(je.n);
        const T *x = reinterpret_cast<const T *>(je.src ? s1 : s0) + (size_t)je.tr * L;
        T *y = reinterpret_cast<T *>(je.src ? s0 : s1) + (size_t)je.tr * L;
        T xr[PER], xi[PER];
        double ld[PER], lm[PER], lp[PER];    // L_kk, Loff[k - 1], Loff[k]
#pragma unroll
        for (int m = 0; m < PER; m++) {
            const int k = tid + m * kJumpThreads;
            xr[m] = xi[m] = T(0);
            ld[m] = lm[m] = lp[m] = 0.0;
            if (k < d) {
                xr[m] = x[k];
                xi[m] = x[dp + k];
                ld[m] = Ldiag[k];
                if (k > 0) lm[m] = Loff[k - 1];
                if (k + 1 < d) lp[m] = Loff[k];
            }
        }
#pragma unroll
        for (int m = 0; m < PER; m++) {
            const int k = tid + m * kJumpThreads;
            if (k < d) {
                pre[k + 1] = (double)xr[m] / sn;
                pim[k + 1] = (double)xi[m] / sn;
            }
        }
        if (tid == 0) {
            pre[0] = pim[0] = 0.0;
            pre[d + 1] = pim[d + 1] = 0.0;
        }
        __syncthreads();
        double vre[PER], vim[PER];
        double p = 0.0;
#pragma unroll
        for (int m = 0; m < PER; m++) {
            const int k = tid + m * kJumpThreads;
            vre[m] = vim[m] = 0.0;
            if (k < d) {
                double rr = 0.0, ii = 0.0;
                if (k > 0) {
                    const double l = -lm[m];
                    rr += l * pre[k];
                    ii += l * pim[k];
                }
                rr += ld[m] * pre[k + 1];
                ii += ld[m] * pim[k + 1];
                if (k + 1 < d) {
                    const double l = lp[m];
                    rr += l * pre[k + 2];
                    ii += l * pim[k + 2];
                }
                vre[m] = rr;
                vim[m] = ii;
                p += rr * rr + ii * ii;
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
        for (int m = 0; m < PER; m++) {
            const int k = tid + m * kJumpThreads;
            if (k < d) {
                y[k] = (T)(vre[m] / sp);
                y[dp + k] = (T)(vim[m] / sp);
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// seed: every trajectory sits at a top-level step boundary between runs, so it
// starts at level 0 of its current branch; target = pos + n.
// ---------------------------------------------------------------------------
__global__ void seed_zero_kernel(int *cnt)
{
    if (threadIdx.x < kMaxGroups) cnt[threadIdx.x] = 0;
}

__global__ void seed_kernel(TrajState *st, int M, int n, int *list, int *cnt, int D)
{
    const int tr = blockIdx.x * blockDim.x + threadIdx.x;
    if (tr >= M) return;
    TrajState s = st[tr];
    s.target = s.pos + n;
    st[tr] = s;
    if (n > 0) {
        const int g = s.b * D;
        const int q = atomicAdd(&cnt[g], 1);
        list[(size_t)g * M + q] = (tr << 1) | s.cur;
    }
}

void launch_seed(TrajState *st, int M, int n_steps, int *list, int *cnt, int D, hipStream_t s)
{
    seed_zero_kernel<<<1, kMaxGroups, 0, s>>>(cnt);
    seed_kernel<<<(M + 255) / 256, 256, 0, s>>>(st, M, n_steps, list, cnt, D);
}

void launch_plan(const RunArgs &a, hipStream_t s)
{
    plan_kernel<<<1, kPlanThreads, 0, s>>>(a);
}

void launch_gemm(int prec, int gemm_kind, const RunArgs &a, const StepGrid &g, hipStream_t s)
{
    if (gemm_kind == kGemmFma) {
        if (prec == kFP64) gemm_fma<double><<<g.gemm, 256, 0, s>>>(a);
        else gemm_fma<float><<<g.gemm, 256, 0, s>>>(a);
    } else if (prec == kFP64) {
        launch_gemm_f64(a, g, s);
    } else if (prec == kBF16x3) {
        if (a.bn == 256) gemm_cp<kCpBF16x3, 256, 4><<<g.gemm, 512, 0, s>>>(a);
        else if (g.waves == 8) gemm_cp<kCpBF16x3, 128, 4><<<g.gemm, 512, 0, s>>>(a);
        else if (a.bn == 64) gemm_cp<kCpBF16x3, 64, 2><<<g.gemm, 256, 0, s>>>(a);
        else gemm_cp<kCpBF16x3, 128, 2><<<g.gemm, 256, 0, s>>>(a);
    } else {
        if (a.bn == 256) gemm_cp<kCpF32x16, 256, 4><<<g.gemm, 512, 0, s>>>(a);
        else if (a.bn == 64) gemm_cp<kCpF32x16, 64, 2><<<g.gemm, 256, 0, s>>>(a);
        else gemm_cp<kCpF32x16, 128, 2><<<g.gemm, 256, 0, s>>>(a);
    }
}

void launch_tail(int prec, const RunArgs &a, const StepGrid &g, hipStream_t s)
{
    decide_kernel<<<g.decide, 256, 0, s>>>(a);
    const size_t lds = 2 * (size_t)(g.jdim + 2) * sizeof(double);
    if (g.jdim <= 4 * kJumpThreads) {
        if (prec == kFP64) jump_kernel<double, 4><<<g.jump, kJumpThreads, lds, s>>>(a, g.jdim);
        else jump_kernel<float, 4><<<g.jump, kJumpThreads, lds, s>>>(a, g.jdim);
    } else {
        static_assert(kMaxDim <= 8 * kJumpThreads, "jump layout");
        if (prec == kFP64) jump_kernel<double, 8><<<g.jump, kJumpThreads, lds, s>>>(a, g.jdim);
        else jump_kernel<float, 8><<<g.jump, kJumpThreads, lds, s>>>(a, g.jdim);
    }
}

int gemm_blocks_per_cu(int prec, int gemm_kind, int bn, int waves)
{
    int n = 0;
    hipError_t e;
    const bool narrow = bn == 64;
    if (gemm_kind == kGemmFma)
        e = prec == kFP64 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, gemm_fma<double>, 256, 0)
                          : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, gemm_fma<float>, 256, 0);
    else if (prec == kFP64)
        e = gemm_f64_occupancy(&n, bn);
    else if (prec == kBF16x3)
        e = bn == 256 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, gemm_cp<kCpBF16x3, 256, 4>, 512, 0)
            : waves == 8 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, gemm_cp<kCpBF16x3, 128, 4>, 512, 0)
            : narrow   ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, gemm_cp<kCpBF16x3, 64, 2>, 256, 0)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, gemm_cp<kCpBF16x3, 128, 2>, 256, 0);
    else
        e = bn == 256 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, gemm_cp<kCpF32x16, 256, 4>, 512, 0)
            : narrow  ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, gemm_cp<kCpF32x16, 64, 2>, 256, 0)
                      : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, gemm_cp<kCpF32x16, 128, 2>, 256, 0);
    return (e == hipSuccess && n > 0) ? n : 1;
}

// ---------------------------------------------------------------------------
// Lane-layout probe of the three matrix-core builtins (one wave, row-major I/O).
// ---------------------------------------------------------------------------
__global__ void mfma_probe_kernel(const float *A32, const float *B32, float *D32, const double *A64,
                                  const double *B64, double *D64, const float *Abf, const float *Bbf,
                                  float *Dbf)
{
    const int l = threadIdx.x;
    floatx16 c = {};
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(A32[(l % 32) * 2 + l / 32], B32[(l / 32) * 32 + l % 32], c, 0, 0, 0);
    for (int r = 0; r < 16; r++) D32[(8 * (r / 4) + 4 * (l / 32) + r % 4) * 32 + l % 32] = c[r];
    doublex4 e = {0.0, 0.0, 0.0, 0.0};
    e = __builtin_amdgcn_mfma_f64_16x16x4f64(A64[(l % 16) * 4 + l / 16], B64[(l / 16) * 16 + l % 16], e, 0, 0, 0);
    for (int r = 0; r < 4; r++) D64[(l / 16 + 4 * r) * 16 + l % 16] = e[r];
    bf16x8 av, bv;
    for (int q = 0; q < 8; q++) {
        av[q] = (__bf16)Abf[(l % 32) * 16 + 8 * (l / 32) + q];
        bv[q] = (__bf16)Bbf[(8 * (l / 32) + q) * 32 + l % 32];
    }
    floatx16 cb = {};
    cb = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, cb, 0, 0, 0);
    for (int r = 0; r < 16; r++) Dbf[(8 * (r / 4) + 4 * (l / 32) + r % 4) * 32 + l % 32] = cb[r];
}

// v_mfma_f32_16x16x4f32: A (16 x 4): lane l holds A[l % 16][l / 16]; B (4 x 16):
// B[l / 16][l % 16]; D (16 x 16): lane l, register r holds D[4 (l / 16) + r][l % 16]
__global__ void mfma_probe16_kernel(const float *A, const float *B, float *D)
{
    const int l = threadIdx.x;
    floatx4 c = {0.f, 0.f, 0.f, 0.f};
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(A[(l % 16) * 4 + l / 16], B[(l / 16) * 16 + l % 16], c, 0, 0, 0);
    for (int r = 0; r < 4; r++) D[(4 * (l / 16) + r) * 16 + l % 16] = c[r];
}

void launch_mfma_probe16(const float *A, const float *B, float *D, hipStream_t s)
{
    mfma_probe16_kernel<<<1, 64, 0, s>>>(A, B, D);
}

void launch_mfma_probe(const float *A32, const float *B32, float *D32, const double *A64,
                       const double *B64, double *D64, const float *Abf, const float *Bbf, float *Dbf,
                       hipStream_t s)
{
    mfma_probe_kernel<<<1, 64, 0, s>>>(A32, B32, D32, A64, B64, D64, Abf, Bbf, Dbf);
}

#endif  // !QJX_STEP_F64
}  // namespace qjx
