// qjx_ctx.h -- the engine context (one dimer model + its ensemble), shared by the
// engine and the job pool.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/qjx.h"
#include "qjx_internal.h"
#include "qjx_runner.h"

struct qjx_ctx {
    qjx_model_desc model{};
    int d = 0, dp = 0, L = 0, D = 0, S = 0, G = 0, R = 0;
    int prec = 0;
    int gemm_kind = 0;
    size_t esize = 8;              // state element bytes
    size_t pesize = 8;             // propagator element bytes (per half for BF16x3)
    double T = 0.0;
    std::vector<double> dt;
    hipStream_t stream = nullptr;

    // model (host)
    std::vector<double> Hdiag, Hoff, Hdrv, Ldiag, Loff, HdiagE;
    std::vector<double> Hqj[2];    // dense interleaved complex d x d

    // device model
    void *props = nullptr;         // 2*D * prop_elems elements
    void *props_lo = nullptr;      // BF16x3: low halves
    bool cp = false;               // complex-pair layout (dp x L per propagator), else block-real (L x L)
    size_t prop_elems = 0;         // elements per propagator
    int *krange = nullptr;         // [G][R][2] K-block ranges of the propagators' row tiles
    unsigned *kfrag = nullptr;     // [G][R][2] the same per 32-row fragment (one byte each)
    bool dense = false;            // QJX_DENSE=1: every K block (no band table)
    double *d_dt = nullptr, *d_Ldiag = nullptr, *d_Loff = nullptr;
    double *d_HdiagE = nullptr, *d_Hoff = nullptr;

    // ensemble
    int M = 0;
    qjx_ensemble_desc ens{};
    uint64_t stride = 0;
    void *slot[2] = {nullptr, nullptr};
    qjx::TrajState *st = nullptr;
    int *lists[2] = {nullptr, nullptr};
    int *cnts[2] = {nullptr, nullptr};
    double *normpart = nullptr;
    int *err = nullptr;
    unsigned long long *counters = nullptr;
    qjx::JumpLog log{};
    double *mean_start = nullptr, *obs = nullptr, *adr = nullptr, *adr_avg = nullptr;
    double *xfer = nullptr;        // export/import buffer (M x d complex)
    bool have_ens = false;
    qjx::Ens desc{};               // device descriptor of this ensemble
    qjx::Runner *runner = nullptr; // drives iterations over {this ensemble}

    // host jump logs
    std::vector<std::vector<double>> jt, jn, je;
    std::vector<std::vector<int>> jty;

    bool profiling = false;        // per-kernel timing of the iterations

    // random observables of the STD path (qjx_random_obs): the matrices of the last
    // (num, seed, mns, type) and scratch for their evaluation
    double2 *robs_R = nullptr;
    int robs_num = 0, robs_seed = 0, robs_mns = 0, robs_type = 0;
    double2 *robs_scratch = nullptr;
    size_t robs_scratch_n = 0;

    // Runge-Kutta propagation (prop_id 1, qjx_ctx_create_rk): no propagator tree;
    // every trajectory takes rk_ns RK4 steps per period in the RK kernel.
    int prop = 0;                  // 0 quantum-jump tree, 1 Runge-Kutta
    int rk_ns = 0, rk_drv_type = 0;
    double rk_phase = 0.0, rk_step = 0.0;
    long long rk_pos = 0;          // RK steps done since period 0 of the current phase
    double *d_rkbands = nullptr;   // band coefficients, lane-major (qjx_rk.h)
    double *d_rktab = nullptr;     // drive table of one launch (3 doubles per step)
    long long rk_tab_steps = 0;    // steps d_rktab holds
    long long rk_launches = 0;
    double rk_ms = 0.0;            // device time of the RK launches (profiling on)
};

namespace qjx {
// Internal entry points used by the job pool.
int ctx_create_kind(const qjx_model_desc *m, int gemm_kind, qjx_ctx **out);
size_t ensemble_footprint(const qjx_model_desc *m, int M, int gemm_kind);
size_t model_footprint(const qjx_model_desc *m, int gemm_kind);
int ctx_seed(qjx_ctx *c, int n_steps, int is_obs);   // draw-free seeding of a run on c->stream
}  // namespace qjx
