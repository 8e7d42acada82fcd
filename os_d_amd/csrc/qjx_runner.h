// qjx_runner.h -- host driver of the propagation iterations over a table of
// ensembles (one or many models), with optional HIP-graph capture.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "qjx_internal.h"
#include "qjx_kernels.h"

namespace qjx {

struct Runner {
    int prec = 0, gemm_kind = 0;
    hipStream_t stream = nullptr;     // not owned
    Ens *d_ens = nullptr;             // [kMaxEns]
    int *d_nens = nullptr;
    Plan *d_plan = nullptr;
    JumpEnt *d_jl = nullptr;          // [capacity]
    int *d_jcnt = nullptr;
    int *h_total = nullptr;           // pinned
    int capacity = 0;                 // trajectories the table may hold in total
    StepGrid grid{};
    int cus = 0;                      // compute units of the device
    int bn = 128;                     // GEMM column tile width (64 or 128)
    int wg_per_cu = 1;                // persistent GEMM workgroups per compute unit
    int sc = 1;                       // column tiles per super-column of the tile order
    int rc = 1;                       // row tiles per GEMM work unit (complex-pair GEMM)
    bool tune_rc = false;             // the tuner also picks rc (1 or 2)
    bool autotune = true;             // pick (bn, wg_per_cu) by timing the live table
    bool tuned = false;
    bool cached = false;              // the shape came from the process-wide cache
    int tunes = 0;                    // shapes this runner has measured
    int nens = 0;
    std::vector<Ens> h_ens;           // host copy of the table
    bool graphs = true;               // replay a captured pair of iterations
    bool defer_capture = false;       // capture nothing now (another thread of the owner
                                      // may call the runtime): launch eagerly meanwhile
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool profiling = false;           // time each kernel with events (no graphs)
    std::vector<hipEvent_t> ev;
    double plan_ms = 0.0, gemm_ms = 0.0, tail_ms = 0.0;
    long long gemm_launches = 0, iterations = 0;
};

int runner_create(int prec, int gemm_kind, hipStream_t stream, int capacity, Runner **out);
void runner_destroy(Runner *r);
// Copy the ensemble table (n descriptors) to the device (on r->stream).
int runner_set_table(Runner *r, const Ens *ens, int n);
// Launch `iters` iterations (even: the lists read next are parity 0 again).
int runner_chunk(Runner *r, int iters);
// Entries the next iteration will process (synchronises r->stream).
int runner_next_total(Runner *r, int *total);
// Chunks of growing length until no trajectory in the table has steps left.
int runner_run(Runner *r, long long *iterations);
// Launch-shape tuner: pick the GEMM column tile width and the persistent grid
// for the current table, from a process-wide cache or by timing the candidates
// on the table's own iterations (they are real steps; results do not depend on
// the shape).  No-op once tuned, for the FMA kernel, or with autotune off; with
// `again`, a runner that measured its shape measures again (the table's mix of
// groups has changed since) and updates the cache.
int runner_autotune(Runner *r, bool again = false);
// Use tile width bn, wg_per_cu workgroups per CU, `waves` waves per GEMM workgroup
// and (rc > 0) rc row tiles per work unit (drops a captured graph when any changes).
void runner_apply_shape(Runner *r, int bn, int wg_per_cu, int waves, int rc = 0);

void set_last_error(const std::string &msg);

}  // namespace qjx
