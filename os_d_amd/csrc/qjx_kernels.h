// qjx_kernels.h -- host-callable launchers of the HIP kernels (qjx_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qjx_internal.h"

namespace qjx {

constexpr int kGemmFma = 0;      // LDS-tiled vector FMA, 64 x 64 tiles
constexpr int kGemmMfma = 1;     // matrix cores, 128 x 128 tiles
// rows of the block-real state per GEMM row tile
inline int gemm_bm(int kind) { return kind == kGemmMfma ? 128 : 64; }
// Whether the propagators are stored as complex-pair rows [Re U | Im U] (dp x 2dp)
// for the complex-pair GEMM (the FP32 and 3xBF16 forms on the matrix cores); the
// FP64 form and the vector-FMA GEMM use the block-real operand.
bool cp_layout(int prec, int gemm_kind);

// Per-trajectory initial state: psi = e_{start_state} (start_type 0) or the
// uniform vector (otherwise); RNG state x_{seed+tr} of the MCG59 leap-frog stream.
void launch_init_ensemble(int prec, void *slot0, void *slot1, TrajState *st, int M, int d, int dp,
                          int L, int seed, int start_type, int start_state, hipStream_t s);
// eta = uniform(); redraw while 0 (experiment.cpp:3259-3265)
void launch_draw_eta(TrajState *st, int M, uint64_t stride, hipStream_t s);
// target = pos + n_steps for every trajectory; list/cnt (one parity of the
// ensemble's group lists) then hold every trajectory at level 0 of its branch.
void launch_seed(TrajState *st, int M, int n_steps, int *list, int *cnt, int D, hipStream_t s);
void launch_reset_times(TrajState *st, int M, hipStream_t s);
void launch_reset_logn(TrajState *st, int M, hipStream_t s);

// Persistent grid sizes of the iteration kernels.
struct StepGrid {
    int gemm;     // workgroups of the GEMM (multiple of 8: XCD-aware tile order)
    int decide;   // workgroups of 256 threads
    int jump;     // workgroups of 256 threads
    int waves = 4; // waves per GEMM workgroup (8: the 3xBF16 kernel's 2 x 4 wave layout)
    int jdim = kMaxDim; // largest Hilbert dimension of the table (jump kernel LDS)
};
// The three launches of one propagation iteration over the ensembles of a.ens:
// plan, grouped GEMM, then decide (crossing test, bisection bookkeeping, next
// lists) and jumps.  The caller checks each launch.
void launch_plan(const RunArgs &a, hipStream_t s);
void launch_gemm(int prec, int gemm_kind, const RunArgs &a, const StepGrid &g, hipStream_t s);
void launch_tail(int prec, const RunArgs &a, const StepGrid &g, hipStream_t s);
// resident GEMM workgroups per compute unit (occupancy of the kernel for the
// column tile width bn and `waves` waves per workgroup)
int gemm_blocks_per_cu(int prec, int gemm_kind, int bn, int waves);
// one wave: D32 = A32 (32x2) B32 (2x32), D64 = A64 (16x4) B64 (4x16),
// Dbf = Abf (32x16) Bbf (16x32) through bf16 operands, all row-major
void launch_mfma_probe(const float *A32, const float *B32, float *D32, const double *A64,
                       const double *B64, double *D64, const float *Abf, const float *Bbf, float *Dbf,
                       hipStream_t s);

// one wave: D = A (16x4) B (4x16) on v_mfma_f32_16x16x4f32, row-major
void launch_mfma_probe16(const float *A, const float *B, float *D, hipStream_t s);

// Observables of every trajectory (core.cpp:665-725): norm, mean, m2, energy;
// dispersion against mean_start; optional |psi_k|^2 / norm per trajectory.
void launch_chars(int prec, const void *slot0, const void *slot1, const TrajState *st, int M,
                  int d, int dp, int L, const double *Hdiag, const double *Hoff, int is_start,
                  double *mean_start, double *norm, double *mean, double *disp, double *m2,
                  double *energy, double *adr, hipStream_t s);
// adr_avg[k] = sum_tr adr[tr][k] / M  (dump.cpp:26-61), fixed summation order.
void launch_adr_avg(const double *adr, int M, int d, double *out, hipStream_t s);
// Copy the current state of every trajectory into a dense interleaved complex
// FP64 array (M x d).
void launch_export_state(int prec, const void *slot0, const void *slot1, const TrajState *st,
                         int M, int d, int dp, int L, double *out, hipStream_t s);
void launch_import_state(int prec, void *slot0, void *slot1, const TrajState *st, int M, int d,
                         int dp, int L, const double *in, hipStream_t s);

// Ensemble density matrix rho = (1/M) sum psi psi^+ / ||psi||^2 (d x d complex
// FP64, row-major interleaved) and purity Tr rho^2.  Scratch: w[M] and
// part[splits][L][L] doubles, splits = density_splits(M, L).
int density_splits(int M, int L);
// out[0] = sum |rho_e|^2 over n complex entries (fixed order, one workgroup)
void launch_purity(const double *rho, size_t n, double *out, hipStream_t s);
void launch_density(int prec, const void *slot0, const void *slot1, const TrajState *st, int M, int d,
                    int dp, int L, double *w, double *part, int splits, double *rho, double *purity,
                    hipStream_t s);

// K-block ranges [lo, hi) per 128-row tile of a packed operand (out: [L/128][2]) and
// per 32-row fragment (frag: byte f of word 0 / 1 = fragment f's lo / hi); dense =
// every block.
// Complex-pair layout (cp): the same per 64-row tile of P (dp x L) in K blocks of 16
// complex columns (out: [dp/64][2]); 3xBF16: two 32-row fragments per tile (bytes 0
// and 1 of frag words 0 and 1); FP32: four 16-row fragments, word f = window
// [lo, hi) in units of 4 complex columns, lo | hi << 16.  frag has kFragWords words
// per tile.
void launch_band(int prec, bool cp, const void *P, const void *Plo, int L, int dp, bool dense, int *out,
                 unsigned *frag, hipStream_t s);

// Dense complex FP64 helpers for the propagator build (row-major n x n, interleaved).
void launch_zgemm(int n, const double *A, const double *B, double *C, hipStream_t s);
// C = alpha * A + beta * I  (alpha, beta real)
void launch_zaxpby_eye(int n, const double *A, double alpha, double beta, double *C, hipStream_t s);
// C += alpha * A
void launch_zaxpy(int n, const double *A, double alpha, double *C, hipStream_t s);
// Write U (n x n complex FP64) as the block-real 2dp x 2dp operand, or (cp) as
// complex-pair rows [Re U_i,: | Im U_i,:] (dp x 2dp).
void launch_pack_propagator(int prec, bool cp, int n, int dp, const double *U, void *dst, void *dst_lo,
                            hipStream_t s);

}  // namespace qjx
