// qjx_obs.h -- observables evaluated on device-resident states in the reference's
// own operation order: random observables (get_random_obs, experiment.cpp:2023-2078)
// and the LPN task's per-period bookkeeping (calc_chars_std / calc_chars_lpn,
// calc_delta_s / calc_delta_f, lambda_lpn, var_trajectory_lpn; core.cpp:665-915,
// experiment.cpp:2221-2393, 3065-3134).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qjx_internal.h"

namespace qjx {

// C = A B for complex matrices (interleaved double2), A rows x K, B K x cols, all
// row-major with leading dimensions lda, ldb, ldc.  Every element accumulates over
// k in ascending order as re = re + (ar br - ai bi), im = im + (ar bi + ai br), with
// no contraction: the loop of the reference's zgemv / the random observables.
void launch_zgemm_seqk(int rows, int cols, int K, const double2 *A, int lda, const double2 *B, int ldb, double2 *C,
                       int ldc, hipStream_t s);
// Scale every entry of the d x d matrix R by 1 / (Tr R) (type 1) or 1 / (Tr R / d)
// (type 2), the trace summed in ascending order (init_random_obs).
void launch_robs_normalise(double2 *R, int d, int type, hipStream_t s);

// Random observables <phi|R_o|phi> / ||phi||^2 of trajectories [tr0, tr0 + n) of an
// ensemble (state slots as in qjx_internal.h), num observables: out[tr][o].
// pn (d x n double2) and y (d x n double2) are scratch.
void launch_random_obs(int prec, const void *slot0, const void *slot1, const TrajState *st, int tr0, int n, int d,
                       int dp, int L, const double2 *R, int num, double2 *pn, double2 *y, double2 *out,
                       hipStream_t s);

// MCG31 stream leap-frog(k, n) of seed: first state x and stride a^n.
void mcg31_stream(uint64_t seed, int k, int n, uint64_t *x, uint64_t *stride);
// The random observables seed + o (o < num) of init_random_obs as device matrices
// [num][d][d] (allocated here; the caller frees).
int build_random_obs(int d, int num, int seed, int mns, int type, double2 **out, hipStream_t s);

// Device side of the LPN bookkeeping.
struct LpnArgs {
    int M, half, d, dp, L, prec, num;        // num random observables
    void *slot0, *slot1;
    TrajState *st;
    const double *Hdiag, *Hoff, *Hdrv;
    double E0, max_energy, T;
    const double2 *R;                        // [num][d][d]
    int lpn_type, eps_deep, lambda_per_periods, save;
    double eps_error, eps_high, eps_low, delta_s_target, df_high, df_low;
    // per trajectory
    double *norm, *mean, *disp, *m2, *energy, *mean_start;
    double *mean_lpn, *energy_lpn, *lambda, *lambda_now, *delta_s;
    int *num_renorms;
    double2 *obs, *obs_lpn;                  // [M][num]
    double *adr;                             // [M][d]
    uint64_t *var_x;                         // perturbation streams (MCG31), one per trajectory
    uint64_t var_stride;
    double2 *work;                           // [M][2 d]: perturbation direction + trial state
    // renormalisations of the last period (save_lambdas): flag, df / ds, ds, df
    int *rec_flag;
    double *rec_ratio, *rec_ds, *rec_df;
};

// calc_chars_std (is_start: calc_chars_std_start) of trajectories [tr0, tr0 + n).
void launch_lpn_chars_std(const LpnArgs &a, int tr0, int n, int is_start, hipStream_t s);
// calc_chars_lpn of trajectories [tr0, tr0 + n).
void launch_lpn_chars_lpn(const LpnArgs &a, int tr0, int n, hipStream_t s);
// trans_process's pairing: every partner copies its original's state, RNG state and
// eta, is perturbed to distance lpn_delta_s, then calc_chars_std_start and
// calc_chars_lpn_start (originals first: launch_lpn_start on [0, half)).
void launch_lpn_pair_start(const LpnArgs &a, hipStream_t s);
void launch_lpn_start(const LpnArgs &a, int tr0, int n, hipStream_t s);
// One period's pair step after the period's calc_chars_std (all) and calc_chars_lpn
// (originals): distance, renormalisation (lambda_lpn) or lambda_lpn_now.
void launch_lpn_pairs(const LpnArgs &a, double curr_time, hipStream_t s);

}  // namespace qjx
