// qjx_internal.h -- device data layout shared by the HIP kernels and the host engine.
//
// HBM layout of one ensemble (d = N + 1 basis states, M trajectories):
//   * states: two slots of M columns, each column 2*dp reals
//     [Re psi_0 .. Re psi_{dp-1}, Im psi_0 .. Im psi_{dp-1}] (dp = d rounded up to the
//     tile multiple; the padding stays zero).  TrajState.cur says which slot holds the
//     current state; a propagation step reads that slot and writes the other one, so
//     accepting a step is a bit flip and rejecting it (bisection descent) costs nothing.
//   * propagators: 2 branches x D levels of U = exp(-i H_eff,b dt_i), row-major, so a
//     row segment of U and a column segment of the state are both contiguous.  The
//     FP64 form (and the vector-FMA GEMM) stores the real block form
//     [[Re U, -Im U], [Im U, Re U]] (2dp x 2dp); the FP32 and 3xBF16 forms on the
//     matrix cores store complex-pair rows [Re U_i,: | Im U_i,:] (dp x 2dp), which hold
//     every entry once (complex-pair GEMM tiles, qjx_step.hip).
//   * per-trajectory scalar state (TrajState), group lists for the grouped GEMM and
//     per-row-tile partial norms.
#pragma once
#include <stdint.h>

namespace qjx {

enum Precision { kFP64 = 0, kFP32 = 1, kBF16x3 = 2 };

// Scalar state of one trajectory: where it is in the bisection tree and its
// FP64 bookkeeping (time, eta, RNG).  Mirrors one_period_branch's recursion
// (reference qj_proc.cpp:271-343): `lvl` is the recursion depth, `cnt0` the
// counter of the top level, bit i of `bits` the counter of level i >= 1.
struct TrajState {
    double t;        // time (times_all)
    double eta;      // jump threshold (etas_all)
    uint64_t rng;    // MCG59 leap-frog state x (next output is x * 2^-59)
    int cnt0;        // steps done at level 0 of the current half period
    uint32_t bits;   // counters of levels 1..D-1
    int lvl;         // current level
    int b;           // branch (half period) 0 / 1
    int pos;         // top-level steps completed (period = pos / S)
    int target;      // run until pos == target
    int cur;         // slot holding the current state
    int njumps;      // jumps logged (jumps_counts)
    int logn;        // entries in the device jump log since the last drain
    int pad;
};

struct JumpLog {
    double *time;    // [M * cap]
    double *norm;
    double *eta;
    int *type;
    int cap;         // records per trajectory between drains (0 = logging off)
};

constexpr int kMaxDim = 2048;                        // largest Hilbert dimension d
constexpr int kMaxGroups = 64;                       // 2 branches x up to 32 levels
constexpr int kMaxEns = 64;                          // ensembles sharing one iteration
constexpr int kPlanItems = kMaxEns * kMaxGroups;     // (ensemble, group) items
// Complex-pair GEMM tiles: 64 complex rows (Re and Im outputs: 128 real rows, as a
// block-real tile) and K blocks of 16 complex columns; band tables of this layout
// count K blocks of 16 and two 32-row fragments per tile.
constexpr int kCpRows = 64;
constexpr int kCpBlock = 16;
// Forms of the complex-pair GEMM: 3xBF16 (32-row fragments, K blocks of 16) and FP32
// (16-row fragments, each with its K window in units of 4 complex columns)
constexpr int kCpBF16x3 = 1, kCpF32x16 = 2;
// band-table words per row tile (kfrag): block-real and 3xBF16 tables use two (one
// byte per fragment in each), the FP32 complex-pair table four (one per 16-row
// fragment: first | past-the-end unit of 4 complex columns << 16)
constexpr int kFragWords = 4;

// Device-resident descriptor of one ensemble (one dimer model and its
// trajectories).  A table of these drives every step kernel, so ensembles of
// different Hilbert dimension share each launch.
struct Ens {
    const void *props;        // 2*D matrices, row-major: L*L elements (L = 2*dp), or dp*L complex-pair
    const void *props_lo;     // BF16x3 only: low halves
    const int *krange;        // [G][R][2]: K blocks (of 32; of kCpBlock complex columns in the
                              // complex-pair layout) [lo, hi) holding every nonzero of row tile
                              // rt within each half of the operand (exact zeros skipped)
    const unsigned *kfrag;    // [G][R][kFragWords]: the same for each 32-row fragment f of the
                              // tile, byte f of word 0 (lo) and of word 1 (hi); FP32 complex-pair:
                              // word f = window of 16-row fragment f in units of 4 columns
    void *slot[2];            // state slots, M*L elements each
    TrajState *st;
    int *list[2];             // per parity: [G * M] entries (trajectory << 1) | current slot
    int *cnt[2];              // per parity: [kMaxGroups] list lengths
    double *normpart;         // [M][R]: FP64 partial sums of y^2 per row tile
    const double *dt;         // [D]
    const double *Ldiag;      // [d]   L_kk
    const double *Loff;       // [d-1] L_{k+1,k} = -Loff[k], L_{k,k+1} = +Loff[k]
    JumpLog log;
    int *err;                 // [4]: code, trajectory, bits of the offending norm
    unsigned long long *counters;  // [0] matvecs, [1] jumps, [2] matrix-core flops the GEMM
                                   // executes, in units of 128
    uint64_t stride;          // MCG59 leap-frog multiplier a^mns
    double norm_tol;          // allowed excess of ||psi||^2 over 1 (0 for FP64)
    int d, dp, L, D, S, M, G, R;
    int is_obs;
    int jump;                 // logging switch (params.txt "jump")
};

// Schedule of one iteration, written by the plan kernel from the list lengths:
// prefixes over items (ensemble e, group g) -> e * kMaxGroups + g.
struct Plan {
    int tile_pre[kPlanItems + 1];   // GEMM tiles (row tiles x column tiles)
    int ent_pre[kPlanItems + 1];    // list entries (trajectories stepping)
    int next_total[2];              // entries appended for the following iteration
    int xnext[8];                   // per-XCD tile claim counters of the GEMM
    unsigned long long cols;        // list entries planned since the runner was created
};

struct JumpEnt {
    int ens, tr, src, pad;    // ensemble, trajectory, slot holding the step result
    double n;                 // its ||psi||^2
};

// Kernel argument of every step kernel (constant for a runner except parity).
struct RunArgs {
    const Ens *ens;           // [kMaxEns]
    const int *nens;          // number of ensembles in the table (device)
    Plan *plan;
    JumpEnt *jl;              // jumps of this iteration
    int *jcnt;
    int parity;               // lists read this iteration: list[parity]
    int bn;                   // GEMM column tile
    int sc;                   // column tiles per super-column (tile order, L2 reuse of U)
    int rc;                   // row tiles per GEMM work unit (complex-pair GEMM; 1 otherwise)
};

}  // namespace qjx
