/*
 * qjx_oracle.h -- FP64 CPU restatement of os_d's QJX quantum-jump path for the
 * driven, dissipative Bose-Hubbard dimer (sys_id 0, task_id 1, prop_id 0).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by the
 * tests, by __graft_entry__.smoke() and by the CPU-baseline leg of bench.py.
 * The product (os_d_amd, libqjx) never links, loads or calls it.
 *
 * Pinning (what makes this restatement trustworthy):
 *   - RNG (qjo_mcg59_*):      bit-exact against golden vectors written by Intel MKL
 *                              2021.4.0 VSL (tests/golden/mcg59_mkl.json).
 *   - expm (qjo_expm):         against scipy.linalg.expm fixtures, the 2x2 closed form,
 *                              the semigroup U(dt)^2 = U(2dt) and unitarity at gamma = 0.
 *   - model (qjo_model_*):     closed forms of H, H_drv, L (hermiticity, L^dag L spectrum
 *                              4 n_-(n_+ + 1) in the c+- basis) and the SU(2) rotation limit.
 *   - stepping / jumps:        brute-force restatements on tiny inputs, norm-monotonicity
 *                              and crossing invariants, gamma = 0 => no jumps, dark state.
 *   - ensemble observables:    the Lindblad rho-space RK4 oracle (lindblad_oracle.h).
 *   - RK path (qjo_rk_*):      the band RHS against the dense -i (H_part + E H_drv) x, RK4
 *                              order-4 convergence to expm, the sub-step closed form with
 *                              the reference's global-rk_step stage arguments, the sine
 *                              against the C library within 2 ulp, the sine drive against
 *                              a DOP853 solution, RNG draw counts, and ensemble averages
 *                              against the Lindblad oracle.
 * No reference file was compiled or imported; the reference needs Eigen and MKL
 * headers that are not present, so it is unbuildable here.
 */
#ifndef QJX_ORACLE_H
#define QJX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Dimer parameters, reference meanings (params.txt keys in brackets). */
typedef struct qjo_model {
    int N;            /* [N] bosons, Hilbert dimension d = N + 1              */
    double prm_E;     /* [dimer_prm_E]                                        */
    double prm_U;     /* [dimer_prm_U]  H_U = 2 (U/N) sum_j n_j (n_j - 1)      */
    double prm_J;     /* [dimer_prm_J]                                        */
    double drv_ampl;  /* [dimer_drv_ampl]  E_b = E +- A on the two half periods */
    double drv_freq;  /* [dimer_drv_freq]  T = 2 pi / omega                   */
    double gamma;     /* [diss_gamma]  rate gamma / N                          */
    int S;            /* top-level steps per period (4 reproduces QJX exactly) */
    int D;            /* [ex_deep] levels of the bisection tree                */
} qjo_model;

/* Dense row-major complex matrices, interleaved (re, im), d*d entries each. */
int qjo_model_matrices(const qjo_model *m, double *H, double *Hdrv, double *L,
                       double *Hqj0, double *Hqj1);

/* exp(A) for a dense row-major complex n x n matrix (Higham 2005 Pade, scaling
 * and squaring).  Returns 0 on success. */
int qjo_expm(int n, const double *A, double *E);

/* MKL-compatible MCG59 leap-frog uniform stream: draws n doubles of stream k. */
void qjo_mcg59_uniform(int k, int mns, int n, double *out);

/* Ensemble handle. */
typedef struct qjo_ens qjo_ens;

qjo_ens *qjo_create(const qjo_model *m, int num_traj, int seed, int mns,
                    int start_type, int start_state, int jump, int num_threads,
                    int build_propagators);
void qjo_destroy(qjo_ens *e);
const char *qjo_error(const qjo_ens *e);

/* Propagator tree access: level lvl of branch b, dense row-major complex d x d. */
int qjo_set_propagator(qjo_ens *e, int b, int lvl, const double *U);
int qjo_get_propagator(const qjo_ens *e, int b, int lvl, double *U);

/* Transient: draw eta_0 per trajectory (redraw while 0), run n periods without
 * logging, then time := 0 (experiment.cpp:411-446, 3249-3271, 2214-2219). */
int qjo_trans(qjo_ens *e, int num_tp_periods);
/* Run n periods for every trajectory, logging jumps when is_obs && jump > 0. */
int qjo_periods(qjo_ens *e, int n, int is_obs);
/* Run n top-level steps (of T/S) for every trajectory; periods are S steps. */
int qjo_steps(qjo_ens *e, int n, int is_obs);
/* Run n periods for one trajectory only (same semantics). */
int qjo_periods_one(qjo_ens *e, int tr, int n, int is_obs);

/* Per-trajectory observables (core.cpp:665-725).  is_start sets mean_start.
 * Any output pointer may be NULL.  adr is num_traj x d. */
int qjo_chars(qjo_ens *e, int is_start, double *norm, double *mean, double *dispersion,
              double *m2, double *energy, double *adr);

/* State read/write.  psi is num_traj x d interleaved complex. */
int qjo_get_state(const qjo_ens *e, double *psi, double *eta, double *t, int *njumps,
                  uint64_t *rng);
int qjo_set_state(qjo_ens *e, const double *psi, const double *eta, const double *t,
                  const uint64_t *rng);
/* Jump log of trajectory tr; returns the number of jumps (copies at most cap). */
int qjo_get_jumps(const qjo_ens *e, int tr, double *times, int *types, double *norms,
                  double *etas, int cap);
/* Counters: matrix-vector products applied (bulk + refinement), jumps applied. */
int64_t qjo_num_matvecs(const qjo_ens *e);
int64_t qjo_num_jumps(const qjo_ens *e);

/* Runge-Kutta propagation (prop_id 1): switch the ensemble to rk_ns RK4 steps
 * per period with drive type drv_type (1: E + A sin(omega t + phase), else the
 * square drive) and phase drv_phase; qjo_trans / qjo_periods / qjo_steps then run
 * RK steps.  Propagators are not needed (create with build_propagators = 0). */
int qjo_rk_setup(qjo_ens *e, int rk_ns, int drv_type, double drv_phase);
/* The sine used by the drive (Cody-Waite reduction + Taylor polynomials). */
double qjo_rk_sin(double x);
/* E(t) of the drive. */
double qjo_rk_drive(const qjo_ens *e, double t);
/* k = -i (H_part + E(t) H_drv) x (x, k: d interleaved complex). */
int qjo_rk_rhs(const qjo_ens *e, double t, const double *x, double *k);
/* ||x||^2 in the chunked summation order of the RK path. */
double qjo_rk_norm(int d, const double *x);
/* One RK4 integration of trajectory tr over `step` from its current time. */
int qjo_rk_int(qjo_ens *e, int tr, double step);

/* MKL VSL MCG31m1 stream leapfrog(k, n) of seed: count uniforms on [lo, hi), or
 * count Box-Muller Gaussians (two uniforms each). */
void qjo_mcg31_uniform(unsigned seed, int k, int n, int count, double lo, double hi, double *out);
void qjo_mcg31_gaussian(unsigned seed, int k, int n, int count, double *out);
/* Random observables seed + o (o < num) of every trajectory's current state
 * (calc_random_obs of the STD path), num_traj x num complex. */
int qjo_random_obs(qjo_ens *e, int num, int seed, int mns, int type, double *out);
/* Random observable obs (d x d complex) of init_random_obs. */
int qjo_random_obs_matrix(int d, int seed, int mns, int type, int obs, double *out);

/* LPN (task_id 7): params.txt keys of the same names. */
typedef struct qjo_lpn_desc {
    int lpn_type;             /* -1 <n2>/d, -2 energy / max|H_ii|, >= 0 random observable */
    int lpn_eps_deep;
    double lpn_eps_error, lpn_eps_high, lpn_eps_low, lpn_delta_s, lpn_delta_f_high, lpn_delta_f_low;
    int lambda_per_periods;   /* 0: renormalise on thresholds, else every period */
    int save_lambdas;
    int num_random_obs, random_obs_seed, random_obs_mns, random_obs_type;
} qjo_lpn_desc;
int qjo_lpn_setup(qjo_ens *e, const qjo_lpn_desc *desc);
int qjo_lpn_trans(qjo_ens *e, int num_tp_periods);
int qjo_lpn_period(qjo_ens *e);
int qjo_lpn_get(qjo_ens *e, double *lambda_now, double *lambda, int *num_renorms, double *mean_lpn,
                double *energy_lpn, double *delta_s, double *mean, double *energy);
int qjo_lpn_random_obs(qjo_ens *e, double *obs, double *obs_lpn);
int qjo_lpn_lambdas(qjo_ens *e, int tr, double *ratio, double *ds, double *df, int cap);
/* 10^y and ln x as the LPN bookkeeping evaluates them (Cody-Waite reduction and
 * Taylor / atanh series, the same code as the library's device side). */
double qjo_lpn_exp10(double y);
double qjo_lpn_log(double x);

#ifdef __cplusplus
}
#endif

#endif
