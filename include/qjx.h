/*
 * qjx.h -- C ABI of the MI355X quantum-jump trajectory engine for the driven,
 * dissipative Bose-Hubbard dimer (os_d QJX, sys_id 0 / task_id 1 / prop_id 0).
 *
 * The reference runs this path as C++ strategy objects over raw arrays, one
 * trajectory per OpenMP iteration (all citations are files of the reference
 * repository, source/cpp/QJX/QJX/):
 *   NewDelBehavior::init_sizes/init_hamiltonians/init_dissipators/init_hamiltonians_qj
 *       (newdel.h:6-17, newdel.cpp:3-37, 450-498, 1454-1498, 2285-2388)  -> qjx_ctx_create
 *   CoreBehavior::init_splits (core.h, core.cpp:6-22) + init_split_branches
 *       (qj_proc.cpp:4-120)                                                -> qjx_ctx_create
 *   ExpNewDelBehavior::init_data (newdel_exp.h:7-12, newdel_exp.cpp:98-117)-> qjx_ensemble_init
 *   ExperimentBehavior::trans_process (experiment.h:8-13, experiment.cpp:411-446)
 *                                                                          -> qjx_trans
 *   PropagateBehavior::one_period (propagator.h:6-23, propagator.cpp:27-30)
 *       over all trajectories, as in obser_process (experiment.cpp:481-494)-> qjx_run_periods
 *   CoreBehavior::calc_chars_std(_start) (core.cpp:665-725)               -> qjx_chars
 *   dump_adr_avg (dump.cpp:26-61)                                          -> qjx_chars(adr_avg)
 *   ed->jump_times/diss_types/jump_norms/jump_etas (data.h, core.cpp:1065-1081)
 *                                                                          -> qjx_get_jumps
 *   main() + Processor::process (main.cpp:3-192, processor.cpp:50-75)     -> qjx_run_config
 * The per-ensemble granularity replaces the per-trajectory calls: a GPU step
 * advances every trajectory at once.
 *
 * Conventions: plain C types only; complex arrays are interleaved (re, im)
 * doubles, row-major; every call returns QJX_OK or a negative code, and
 * qjx_last_error() describes the last failure of the calling thread.  The
 * reference prints and exits (utils.cpp:3-7) or throws (experiment.cpp:1250);
 * this library returns the error instead.  One host thread per context.
 */
#ifndef QJX_H
#define QJX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QJX_OK 0
#define QJX_ERR_ARG (-1)    /* invalid argument / call order                        */
#define QJX_ERR_HIP (-2)    /* HIP runtime failure                                  */
#define QJX_ERR_NORM (-3)   /* ||psi||^2 > 1 after a step (reference: invalid_argument) */
#define QJX_ERR_LOG (-4)    /* jump log overflow between two drains                 */
#define QJX_ERR_IO (-5)     /* config / output file error                           */
#define QJX_ERR_NOMEM (-6)  /* device memory exhausted                              */

#define QJX_FP64 0          /* complex128 states and propagators, FP64 arithmetic   */
#define QJX_FP32 1          /* complex64 states/propagators, FP32 GEMM; FP64 norms  */
#define QJX_BF16X3 2        /* complex64 states, 3-term bf16 split GEMM; FP64 norms  */

typedef struct qjx_ctx qjx_ctx;

/* Dimer model; field names follow params.txt keys, with the reference meanings
 * (dimer_prm_U enters as H_U = 2 (U/N) sum_j n_j (n_j - 1), newdel.cpp:459,484). */
typedef struct qjx_model_desc {
    int N;                  /* N                                                    */
    double prm_E;           /* dimer_prm_E                                          */
    double prm_U;           /* dimer_prm_U                                          */
    double prm_J;           /* dimer_prm_J                                          */
    double drv_ampl;        /* dimer_drv_ampl (square drive E +- A per half period) */
    double drv_freq;        /* dimer_drv_freq (T = 2 pi / freq)                     */
    double diss_gamma;      /* diss_gamma (rate gamma / N)                          */
    int ex_deep;            /* ex_deep: levels of the bisection tree                */
    int steps_per_period;   /* top-level steps per period, even; 4 = reference QJX  */
    int precision;          /* QJX_FP64 / QJX_FP32 / QJX_BF16X3                     */
} qjx_model_desc;

typedef struct qjx_ensemble_desc {
    int num_trajectories;   /* num_trajectories                                     */
    int seed;               /* seed: stream of trajectory tr is leap-frog(seed + tr) */
    int mns;                /* mns: leap-frog stride                                */
    int start_type;         /* start_type: 0 = basis state, else uniform            */
    int start_state;        /* start_state                                          */
    int jump;               /* jump: 0 off, 1 times + types, 2 also norms + etas    */
    int log_capacity;       /* device jump records per trajectory between drains (0 = default) */
} qjx_ensemble_desc;

typedef struct qjx_stats {
    int64_t matvecs;        /* propagator applications (columns of all step GEMMs)  */
    int64_t iterations;     /* grouped step iterations launched                     */
    int64_t jumps;          /* jumps applied                                        */
    double gemm_flops;      /* 8 d^2 per matvec (complex MAC = 8 real flop)         */
    double gemm_ms;         /* device time of the step GEMM launches (profiling on)  */
    double decide_ms;       /* device time of the decide launches (profiling on)     */
    int64_t gemm_launches;  /* step GEMM launches timed                             */
    int32_t gemm_bn;        /* column tile width of the step GEMM (launch-shape tuner) */
    int32_t gemm_grid;      /* persistent GEMM workgroups                           */
    double gemm_band_flops; /* flops of the K blocks the matrix-core GEMM multiplied:
                               2 x 128 x 32 per (column, row tile, K block) inside the
                               propagators' band tables (exact zeros skipped)      */
    int32_t gemm_waves;     /* waves per GEMM workgroup (4 or 8)                     */
    int32_t pad;
} qjx_stats;

const char *qjx_last_error(void);

/* Runge-Kutta propagation (prop_id 1) of the same model. */
typedef struct qjx_rk_desc {
    int rk_ns;              /* rk_ns: RK4 steps per period                          */
    int drv_type;           /* dimer_drv_type: 1 = E + A sin(omega t + phase), any
                               other value the square drive E +- A per half period  */
    double drv_phase;       /* dimer_drv_phase                                      */
} qjx_rk_desc;

/* Build H, H_drv, L, H_eff,b and the propagator tree exp(-i H_eff,b dt_i),
 * dt_0 = T / steps_per_period, dt_i = dt_{i-1} / 2, on the current HIP device. */
int qjx_ctx_create(const qjx_model_desc *model, qjx_ctx **out);
void qjx_ctx_destroy(qjx_ctx *ctx);

/* A context whose trajectories take rk_ns RK4 steps per period instead of the
 * propagator tree: RKPropagateBehavior (propagator.cpp:66-129) with init_rk
 * (rk_proc.cpp:56-63), rk_period (core.cpp:439-450) and rk_step_dimer / rk_int_dimer /
 * rk_right_part_dimer (core.cpp:11033-11115, 11592-11619, 11886-11948).  Every other
 * call below works on it unchanged (ensemble_init, trans, run_periods, run_steps
 * (RK steps), chars, density, state, jumps, stats: matvecs counts right-hand-side
 * evaluations); it has no propagators and is not batched.  model->precision must be
 * QJX_FP64; ex_deep and steps_per_period are ignored.  Arithmetic: each multiply-add
 * one fma, the band of H_part in ascending column order; the stage arguments use the
 * full rk_step for every partial step, as arg_upd does (rk_proc.cpp:141-174). */
int qjx_ctx_create_rk(const qjx_model_desc *model, const qjx_rk_desc *rk, qjx_ctx **out);

/* Allocate the ensemble; states start at e_{start_state} (or uniform), RNG
 * streams at leap-frog(seed + tr, mns) of MCG59 seeded 777. */
int qjx_ensemble_init(qjx_ctx *ctx, const qjx_ensemble_desc *ens);

/* Transient: draw eta_0 (redraw while 0), run num_tp_periods periods without
 * logging, then set every trajectory's time to 0. */
int qjx_trans(qjx_ctx *ctx, int num_tp_periods);

/* Advance every trajectory by n periods; is_obs = 1 logs jumps (if jump > 0). */
int qjx_run_periods(qjx_ctx *ctx, int n, int is_obs);

/* Advance every trajectory by n top-level steps of T / steps_per_period (with
 * all their bisection refinement); n periods = n * steps_per_period steps. */
int qjx_run_steps(qjx_ctx *ctx, int n, int is_obs);

/* Observables of every trajectory (arrays of num_trajectories doubles, any may be
 * NULL); adr_avg (d doubles) is the ensemble mean of |psi_k|^2/||psi||^2;
 * adr (num_trajectories x d) the per-trajectory values.  is_start = 1 also
 * records mean_start (calc_chars_std_start). */
int qjx_chars(qjx_ctx *ctx, int is_start, double *norm, double *mean, double *dispersion,
              double *m2, double *energy, double *adr_avg, double *adr);

/* Ensemble density matrix rho = (1/M) sum_tr |psi_tr><psi_tr| / ||psi_tr||^2
 * (d x d complex, row-major interleaved; may be NULL) and its purity Tr rho^2
 * (may be NULL).  The reference's QJ path outputs only per-trajectory
 * observables; its Lindblad solver writes the final rho of the same model
 * (CQdiss_os_dimer characteristics.cpp:3-6), which this estimates at any dump. */
int qjx_density(qjx_ctx *ctx, double *rho, double *purity);

/* Random observables of the STD path (calc_random_obs, core.cpp:694, 724,
 * 12614-12622): <psi|R_o|psi> / ||psi||^2 for o < num of every trajectory's current
 * state, out = num_trajectories x num complex.  R_o are init_random_obs's matrices
 * (newdel.cpp:3652-3722; random_obs_seed + o, random_obs_mns, random_obs_type), built
 * on the device on first use and kept by the context; the products run in
 * get_random_obs's order (experiment.cpp:2023-2078). */
int qjx_random_obs(qjx_ctx *ctx, int num, int seed, int mns, int type, double *out);

/* qjx_density without device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Its scratch is one fixed 64 MiB
 * device array per GPU, held until the process ends; the call allocates nothing.
 * Limits: d <= 1024 and at most 2^21 trajectories (QJX_ERR_ARG beyond them).  Calls
 * that overlap on one GPU share that array and race: serialise them.  Results equal
 * qjx_density's (bit for bit when its split count fits the scratch, else to rounding). */
int qjx_density_noalloc(qjx_ctx *ctx, double *rho, double *purity);

/* Current states (num_trajectories x d interleaved complex), eta, time, logged
 * jump counts and RNG states; any pointer may be NULL. */
int qjx_get_state(qjx_ctx *ctx, double *psi, double *eta, double *t, int *njumps, uint64_t *rng);
int qjx_set_state(qjx_ctx *ctx, const double *psi, const double *eta, const double *t,
                  const uint64_t *rng);

/* Jump log of trajectory tr accumulated on the host since qjx_ensemble_init.
 * Copies at most cap records, returns the number of records (or < 0). */
int qjx_get_jumps(qjx_ctx *ctx, int tr, double *times, int *types, double *norms, double *etas,
                  int cap);

/* Distribution of the times between consecutive logged jumps of every trajectory,
 * as the reference's post-processing bins the jump_times files
 * (plot/mpipks/qjx/dimer/times_between_jumps.m:41-131): num_decades * bins_per_decade
 * log-spaced bins from bin_begin; pdf (may be NULL) is normalised to unit integral
 * over the in-range intervals, counts (may be NULL) are the raw bin counts, outside
 * (may be NULL) the intervals outside [bin_begin, bin_begin 10^num_decades]. */
int qjx_jump_interval_pdf(qjx_ctx *ctx, double bin_begin, int num_decades, int bins_per_decade,
                          double *pdf, int64_t *counts, int64_t *outside);

/* Propagator of branch b, level lvl as dense d x d complex (FP64 copy of the
 * operand the GEMM uses); set replaces it (e.g. to share one tree between runs). */
int qjx_get_propagator(qjx_ctx *ctx, int b, int lvl, double *U);
int qjx_set_propagator(qjx_ctx *ctx, int b, int lvl, const double *U);

int qjx_get_stats(qjx_ctx *ctx, qjx_stats *out);
/* Time every step kernel with HIP events on the engine stream (resets the timers). */
int qjx_set_profiling(qjx_ctx *ctx, int on);

/* Self-test of the matrix-core operand layouts the GEMM kernels assume (no
 * reference counterpart): D32 = A32 B32 (32x2 by 2x32, FP32), D64 = A64 B64
 * (16x4 by 4x16, FP64) and Dbf = Abf Bbf (32x16 by 16x32 with bf16 operands,
 * FP32 result), all row-major, computed by one wavefront. */
int qjx_mfma_probe(const float *A32, const float *B32, float *D32, const double *A64,
                   const double *B64, double *D64, const float *Abf, const float *Bbf, float *Dbf);
/* The same for the FP32 16x16x4 matrix-core product of the FP32 step GEMM (no
 * reference counterpart): D = A B, A 16x4, B 4x16, D 16x16, row-major, FP32. */
int qjx_mfma_probe_f32x16(const float *A, const float *B, float *D);
/* Measured ceilings of the current device (no reference counterpart): FP32, FP64
 * and BF16 matrix-core rates from back-to-back independent MFMA chains on every
 * compute unit, the FP64 vector FMA rate, and the HBM bandwidth of a 2 GiB device
 * copy (bytes read + written per second); any pointer may be NULL. */
int qjx_peak_probe(double *f32_mfma_tflops, double *f64_mfma_tflops, double *bf16_mfma_tflops,
                   double *f64_valu_tflops, double *hbm_gbps);

/* Hilbert dimension of the context (N + 1). */
int qjx_dim(const qjx_ctx *ctx);

/* ------------------------------------------------------------------------ */
/* Job pool: continuous batching of ensemble jobs (parameter points) on one  */
/* GPU.  The reference runs one ensemble per process and sweeps parameters   */
/* as separate jobs (submit/qjx/submit_dimer.py:128-231); here running jobs  */
/* share every propagation launch whatever their Hilbert dimension, queued   */
/* jobs are admitted while their HBM footprint fits the budget, and a job's  */
/* results do not depend on which other jobs share the launches.             */
/* ------------------------------------------------------------------------ */
/* Batches: several ensembles of one precision and any dimensions advanced
 * together, one grouped GEMM per iteration over all of them.  No counterpart in
 * the reference, which runs one parameter point per process (the job scripts
 * under submit/ sweep parameters by launching processes); a sweep becomes one
 * batch.  The contexts stay owned by the caller and must outlive the batch;
 * each must hold an initialised ensemble.  Results per context are identical
 * to running it alone. */
typedef struct qjx_batch qjx_batch;
int qjx_batch_create(qjx_ctx *const *ctxs, int n, qjx_batch **out);
void qjx_batch_destroy(qjx_batch *batch);
/* n_steps top-level steps of every member (its own T / steps_per_period each). */
int qjx_batch_run_steps(qjx_batch *batch, int n_steps, int is_obs);
/* n_periods periods of every member (n_periods * steps_per_period steps each). */
int qjx_batch_run_periods(qjx_batch *batch, int n_periods, int is_obs);
/* Totals over the members (matvecs, jumps, flops) and the batch's own timings. */
int qjx_batch_get_stats(qjx_batch *batch, qjx_stats *out);
int qjx_batch_set_profiling(qjx_batch *batch, int on);

/* The dimer's Lindblad equation integrated in rho space on the GPU (FP64, RK4,
 * steps_per_period even, the drive constant over each step, steps aligned with the
 * half periods): the density matrix the trajectory ensemble estimates.  Replaces
 * the reference's f-basis integrator CQdiss_os_dimer calcODE_std + calcRho
 * (CQdiss/calcODE.cpp:310-389, CalcRho.cpp:4-67), with rho itself as the state.
 * rho starts as |e_0><e_0|; rho is d x d complex, row-major interleaved. */
typedef struct qjx_lindblad qjx_lindblad;
int qjx_lindblad_create(const qjx_model_desc *model, int steps_per_period, qjx_lindblad **out);
/* The drive of CQdiss_os_dimer's right-hand side (calcVectValue_real,
 * CQdiss/calcODE.cpp:100-116), evaluated at every RK4 stage time t = p T + s h
 * (+ h/2, + h): E + A sin(omega t + phase) for drv_type 1, E + A sign(sin(omega t +
 * phase)) otherwise (a zero sine counts as positive).  per_stage = 0 instead holds the
 * square drive constant over each step, steps aligned with the half periods (what
 * qjx_lindblad_create does; drv_type 0 and phase 0 only).  Time starts at 0 at
 * creation and at every qjx_lindblad_set_rho; each qjx_lindblad_run_periods call
 * continues it. */
typedef struct qjx_lindblad_drive {
    int drv_type;           /* dimer_drv_type                                       */
    double drv_phase;       /* dimer_drv_phase                                      */
    int per_stage;          /* 1: the drive at every stage time (CQdiss); 0: per step */
} qjx_lindblad_drive;
int qjx_lindblad_create_drive(const qjx_model_desc *model, int steps_per_period, const qjx_lindblad_drive *drive,
                              qjx_lindblad **out);
void qjx_lindblad_destroy(qjx_lindblad *lb);
int qjx_lindblad_set_rho(qjx_lindblad *lb, const double *rho);
int qjx_lindblad_get_rho(qjx_lindblad *lb, double *rho);
int qjx_lindblad_run_periods(qjx_lindblad *lb, int n_periods);
/* <k>, <(k - <k>)^2>, Tr(rho (H + E H_drv)), Tr(rho^2) and diag(rho) (d doubles);
 * any pointer may be NULL. */
int qjx_lindblad_observables(qjx_lindblad *lb, double *mean, double *m2, double *energy,
                             double *purity, double *adr);
int qjx_lindblad_dim(const qjx_lindblad *lb);
/* rho written as CQdiss_os_dimer writes its final rho (data.cpp:40-46,
 * save_sparse_complex_mtx utils.cpp:266-310): "row col re im" per entry, 1-based. */
int qjx_lindblad_save_rho(qjx_lindblad *lb, const char *path);

/* ------------------------------------------------------------------------ */
/* LPN task (task_id 7): growth rates of observable distances between paired  */
/* trajectories.  Replaces LpnMultNewDelBehaviour / LpnMultExperimentBehaviour */
/* (newdel_exp.cpp:29-48, experiment.cpp:137-286) with lambda_lpn(_now),       */
/* var_trajectory_lpn, calc_chars_lpn, calc_delta_s/f (experiment.cpp:2221-2393,*/
/* 3065-3134; core.cpp:727-915) and init_random_obs / get_random_obs          */
/* (newdel.cpp:3652-3722, experiment.cpp:2023-2078).  Works on a context of    */
/* either propagation (QJ or RK) with an initialised ensemble of M (even)      */
/* trajectories: tr < M/2 are the originals, tr + M/2 their partners.          */
/* ------------------------------------------------------------------------ */
typedef struct qjx_lpn_desc {
    int lpn_type;           /* lpn_type: -1 |<k> - <k>'| / d, -2 |E - E'| / max|H_ii|,
                               >= 0 |o - o'| of random observable lpn_type          */
    int lpn_eps_deep;       /* lpn_eps_deep: bisection steps of the perturbation size */
    double lpn_eps_error;   /* lpn_eps_error                                        */
    double lpn_eps_high;    /* lpn_eps_high (log10 of the largest perturbation)     */
    double lpn_eps_low;     /* lpn_eps_low                                          */
    double lpn_delta_s;     /* lpn_delta_s: target distance after a renormalisation */
    double lpn_delta_f_high;/* lpn_delta_f_high                                     */
    double lpn_delta_f_low; /* lpn_delta_f_low                                      */
    int lambda_per_periods; /* lambda_per_periods: 0 renormalise on the thresholds, else every period */
    int save_lambdas;       /* save_lambdas: keep every renormalisation's ratio and distances */
    int num_random_obs;     /* num_random_obs                                       */
    int random_obs_seed;    /* random_obs_seed                                      */
    int random_obs_mns;     /* random_obs_mns                                       */
    int random_obs_type;    /* random_obs_type: 0 raw X X^dag, 1 / Tr, 2 / (Tr / d)  */
} qjx_lpn_desc;
typedef struct qjx_lpn qjx_lpn;
int qjx_lpn_create(qjx_ctx *ctx, const qjx_lpn_desc *desc, qjx_lpn **out);
void qjx_lpn_destroy(qjx_lpn *lpn);
/* trans_process: the transient, then every partner copies its original (state,
 * RNG stream, eta) and is perturbed to distance lpn_delta_s; times restart at 0. */
int qjx_lpn_trans(qjx_lpn *lpn, int num_tp_periods);
/* n observation periods (logging jumps as qjx_run_periods with is_obs = 1). */
int qjx_lpn_run_periods(qjx_lpn *lpn, int n);
/* Per trajectory (arrays of M, any may be NULL): lambda_now (the reference's
 * "lambda" output), the accumulated log growth, renormalisations, the lpn mean and
 * energy, the distance after the last renormalisation. */
int qjx_lpn_get(qjx_lpn *lpn, double *lambda_now, double *lambda, int *num_renorms, double *mean_lpn,
                double *energy_lpn, double *delta_s);
/* Random observables (M x num_random_obs complex, interleaved) of the last period
 * (calc_chars_std) and the lpn copies (calc_chars_lpn); either may be NULL. */
int qjx_lpn_random_obs(qjx_lpn *lpn, double *obs, double *obs_lpn);
/* Saved renormalisations of trajectory tr (save_lambdas > 0): delta_f / delta_s,
 * delta_s, delta_f; copies at most cap, returns the count. */
int qjx_lpn_lambdas(qjx_lpn *lpn, int tr, double *ratio, double *ds, double *df, int cap);
/* Random observable obs as d x d complex (interleaved, row-major). */
int qjx_lpn_random_obs_matrix(qjx_lpn *lpn, int obs, double *R);
/* calc_chars_std of the last period, taken before that period's renormalisations
 * (arrays of M), the current abs_diag_rho (M x d, rewritten by calc_chars_lpn as in
 * the reference) and its average over trajectories [begin, end) as dump_adr_avg /
 * dump_adr_avg_mult form it (dump.cpp:26-98); any pointer may be NULL. */
int qjx_lpn_chars(qjx_lpn *lpn, double *norm, double *mean, double *dispersion, double *m2, double *energy,
                  double *adr, int begin, int end, double *adr_avg);

typedef struct qjx_pool qjx_pool;

typedef struct qjx_job_result {
    int status;             /* 0 queued, 1 running, 2 done, < 0 error code          */
    int num_trajectories;
    int n_periods;
    int dim;
    double t_submit;        /* seconds on the pool clock                            */
    double t_admit;
    double t_done;
    int64_t matvecs;        /* propagator applications of this job                  */
    int64_t jumps;
    double mean_avg, mean_std;      /* ensemble mean / std of <k> (= <n2>) at the end */
    double energy_avg, energy_std;  /* ... of the energy                             */
    double norm_avg;
} qjx_job_result;

/* precision applies to every job; hbm_budget_gib <= 0 means 80% of free HBM;
 * max_trajectories bounds the trajectories of all running jobs together. */
int qjx_pool_create(int precision, double hbm_budget_gib, int max_trajectories, qjx_pool **out);
void qjx_pool_destroy(qjx_pool *pool);
/* Queue a job: n_periods periods after eta_0 is drawn (no transient). */
int qjx_pool_submit(qjx_pool *pool, const qjx_model_desc *model, const qjx_ensemble_desc *ens,
                    int n_periods, int *job_id);
/* One scheduling cycle: a chunk of iterations over the running jobs, admission
 * of queued jobs while it runs, retirement of finished jobs. */
int qjx_pool_step(qjx_pool *pool, int *running, int *queued);
/* Status and results; mean / energy (num_trajectories doubles) may be NULL. */
int qjx_pool_result(qjx_pool *pool, int job_id, qjx_job_result *out, double *mean, double *energy);
double qjx_pool_clock(const qjx_pool *pool);
/* Admission state: HBM bytes reserved by admitted jobs (built, being built or
 * running) against the budget, trajectories in the table against the capacity,
 * and job counts by state.  No reference counterpart (the reference has no
 * shared device to admit jobs onto). */
typedef struct qjx_pool_usage {
    int64_t used_bytes;
    int64_t budget_bytes;
    int64_t trajectories;
    int64_t capacity;
    int32_t queued;
    int32_t building;
    int32_t running;
    int32_t done;
    int32_t failed;
    int32_t pad;
} qjx_pool_usage;
int qjx_pool_get_usage(qjx_pool *pool, qjx_pool_usage *out);

/* Drop-in runner: read config.txt / params.txt (reference keys, config.cpp:29-145),
 * run the STD experiment (task_id 1) or the STD deep experiment (task_id 4: a dump
 * after each of the 2 deep_num_steps sub-periods of a period, experiment.cpp:769-862)
 * with QJ (prop_id 0) or RK (prop_id 1) propagation, and write the reference's output
 * files into `path` (core.cpp:918-1084, dump.cpp:26-61).  precision selects the GEMM
 * form of the QJ path; steps_per_period = 0 means the reference's 4 (task 4 fixes it
 * at 8 deep_num_steps: 4 level-0 steps of T / (8 deep_num_steps) per sub-period,
 * init_split_structure_dimer_deep, core.cpp:9928-9981, qj_proc.cpp:15-24, 345-360). */
int qjx_run_config(const char *config_path, const char *params_path, int precision,
                   int steps_per_period);

#ifdef __cplusplus
}
#endif

#endif
