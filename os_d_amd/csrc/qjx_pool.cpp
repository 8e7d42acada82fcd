// qjx_pool.cpp -- continuous batching of ensemble jobs on one GPU.
//
// A job is one dimer parameter point with its trajectory ensemble and a number of
// periods (the reference runs one such ensemble per process, one OpenMP loop per
// period, experiment.cpp:448-542; parameter sweeps are separate jobs,
// submit/qjx/submit_dimer.py).  Here every running job is one entry of a shared
// ensemble table: each propagation iteration is one set of launches over the
// trajectories of all running jobs, whatever their Hilbert dimension.  Jobs are
// admitted while their HBM footprint fits the budget, join the table at the next
// chunk boundary, and retire when their last trajectory has finished; the others
// keep running.  A trajectory's arithmetic does not depend on which other jobs
// share the launches, so a job's results are the same alone and in the mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/qjx.h"
#include "qjx_ctx.h"
#include "qjx_kernels.h"
#include "qjx_runner.h"

namespace {

enum JobState { kQueued = 0, kRunning = 1, kDone = 2, kFailed = 3 };

struct Job {
    int id = 0;
    qjx_model_desc m{};
    qjx_ensemble_desc e{};
    int n_periods = 0;
    int state = kQueued;
    int err = 0;
    std::string msg;
    qjx_ctx *ctx = nullptr;
    size_t footprint = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    std::vector<double> norm, mean, m2, energy;
    long long matvecs = 0, jumps = 0;
};

int pool_fail(int code, const std::string &msg)
{
    qjx::set_last_error(msg);
    return code;
}

}  // namespace

struct qjx_pool {
    int prec = 0;
    int gemm_kind = qjx::kGemmMfma;
    size_t budget = 0, used = 0;
    int capacity = 0, in_table = 0;
    hipStream_t stream = nullptr;
    qjx::Runner *runner = nullptr;
    std::vector<Job *> jobs;           // by id
    std::deque<int> queue;             // waiting for admission
    std::vector<int> running;          // ids in the table order
    bool table_dirty = false;
    long long cycles = 0;                 // scheduling cycles that launched iterations
    // model builds run on a builder thread so that running jobs keep iterating
    std::thread builder;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<int> to_build;          // admitted (budget reserved), not yet built
    std::deque<int> built;             // built and seeded, waiting to join the table
    int building = 0;                  // jobs handed to the builder and not yet in `built`
    bool stop = false;
    int device = 0;
    std::chrono::steady_clock::time_point t0;
    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

extern "C" double qjx_pool_clock(const qjx_pool *p)
{
    return p ? p->now() : 0.0;
}

static int admit(qjx_pool *p, Job *j);

static void builder_main(qjx_pool *p)
{
    (void)hipSetDevice(p->device);
    for (;;) {
        int id;
        Job *j;
        {
            std::unique_lock<std::mutex> lk(p->mu);
            p->cv.wait(lk, [&] { return p->stop || !p->to_build.empty(); });
            if (p->stop) return;
            id = p->to_build.front();
            p->to_build.pop_front();
            j = p->jobs[id];          // `jobs` grows under the same lock
        }
        const int rc = admit(p, j);
        if (rc) {
            j->err = rc;
            j->msg = qjx_last_error();
            if (j->ctx) qjx_ctx_destroy(j->ctx);
            j->ctx = nullptr;
        }
        std::lock_guard<std::mutex> lk(p->mu);
        p->built.push_back(id);
    }
}

extern "C" int qjx_pool_create(int precision, double hbm_budget_gib, int max_trajectories, qjx_pool **out)
{
    if (!out) return pool_fail(QJX_ERR_ARG, "null argument");
    *out = nullptr;
    if (precision != QJX_FP64 && precision != QJX_FP32 && precision != QJX_BF16X3)
        return pool_fail(QJX_ERR_ARG, "unknown precision form");
    if (max_trajectories < 1) return pool_fail(QJX_ERR_ARG, "max_trajectories must be >= 1");
    qjx_pool *p = new qjx_pool;
    p->prec = precision;
    p->capacity = max_trajectories;
    p->t0 = std::chrono::steady_clock::now();
    size_t freeb = 0, totalb = 0;
    if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) {
        delete p;
        return pool_fail(QJX_ERR_HIP, "no HIP device available");
    }
    p->budget = hbm_budget_gib > 0 ? (size_t)(hbm_budget_gib * (1ULL << 30)) : (size_t)(0.8 * (double)freeb);
    // Non-blocking: the builder thread's null-stream calls (hipMemcpy, hipMemset,
    // hipMalloc) then neither wait for the running chunk nor invalidate a graph
    // capture in progress on this stream.  Every hand-over between the two is
    // ordered by the host instead: a build ends with hipStreamSynchronize on the
    // job's own stream, and results are read after synchronising this one.
    if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
        delete p;
        return pool_fail(QJX_ERR_HIP, "hipStreamCreate failed");
    }
    int rc = qjx::runner_create(precision, p->gemm_kind, p->stream, max_trajectories, &p->runner);
    if (rc) {
        (void)hipStreamDestroy(p->stream);
        delete p;
        return rc;
    }
    if (const char *env = std::getenv("QJX_GRAPH"))
        if (std::string(env) == "0") p->runner->graphs = false;
    (void)hipGetDevice(&p->device);
    p->builder = std::thread(builder_main, p);
    *out = p;
    return QJX_OK;
}

extern "C" void qjx_pool_destroy(qjx_pool *p)
{
    if (!p) return;
    {
        std::lock_guard<std::mutex> lk(p->mu);
        p->stop = true;
    }
    p->cv.notify_all();
    if (p->builder.joinable()) p->builder.join();
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    for (Job *j : p->jobs) {
        if (j->ctx) qjx_ctx_destroy(j->ctx);
        delete j;
    }
    qjx::runner_destroy(p->runner);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

extern "C" int qjx_pool_submit(qjx_pool *p, const qjx_model_desc *m, const qjx_ensemble_desc *e, int n_periods,
                               int *job_id)
{
    if (!p || !m || !e || !job_id) return pool_fail(QJX_ERR_ARG, "null argument");
    if (n_periods < 0 || e->num_trajectories < 1) return pool_fail(QJX_ERR_ARG, "bad job size");
    if (e->num_trajectories > p->capacity) return pool_fail(QJX_ERR_ARG, "job larger than the pool capacity");
    Job *j = new Job;
    j->id = (int)p->jobs.size();
    j->m = *m;
    j->m.precision = p->prec;
    j->e = *e;
    j->e.jump = 0;          // serving jobs return observables, not jump logs
    j->n_periods = n_periods;
    j->footprint = qjx::model_footprint(&j->m, p->gemm_kind) +
                   qjx::ensemble_footprint(&j->m, e->num_trajectories, p->gemm_kind);
    if (j->footprint > p->budget) {
        delete j;
        return pool_fail(QJX_ERR_NOMEM, "job footprint exceeds the pool's HBM budget");
    }
    j->t_submit = p->now();
    std::lock_guard<std::mutex> lk(p->mu);
    j->id = (int)p->jobs.size();
    p->jobs.push_back(j);
    p->queue.push_back(j->id);
    *job_id = j->id;
    return QJX_OK;
}

// Build the job's model and ensemble (own stream), draw eta_0 and seed its run
// (builder thread).
static int admit(qjx_pool *p, Job *j)
{
    (void)p;
    int rc = qjx::ctx_create_kind(&j->m, p->gemm_kind, &j->ctx);
    if (rc) return rc;
    rc = qjx_ensemble_init(j->ctx, &j->e);
    if (rc) return rc;
    qjx::launch_draw_eta(j->ctx->st, j->ctx->M, j->ctx->stride, j->ctx->stream);
    rc = qjx::ctx_seed(j->ctx, j->n_periods * j->ctx->S, 1);
    if (rc) return rc;
    if (hipStreamSynchronize(j->ctx->stream) != hipSuccess) return pool_fail(QJX_ERR_HIP, "job setup failed");
    return QJX_OK;
}

static int retire(qjx_pool *p, Job *j)
{
    qjx_ctx *c = j->ctx;
    const int M = c->M;
    j->norm.assign(M, 0.0);
    j->mean.assign(M, 0.0);
    j->m2.assign(M, 0.0);
    j->energy.assign(M, 0.0);
    std::vector<double> disp(M);
    int rc = qjx_chars(c, 0, j->norm.data(), j->mean.data(), disp.data(), j->m2.data(), j->energy.data(),
                       nullptr, nullptr);
    if (rc) return rc;
    qjx_stats st{};
    qjx_get_stats(c, &st);
    j->matvecs = st.matvecs;
    j->jumps = st.jumps;
    int h[4] = {0, 0, 0, 0};
    if (hipMemcpy(h, c->err, sizeof h, hipMemcpyDeviceToHost) == hipSuccess && h[0]) {
        j->state = kFailed;
        j->err = QJX_ERR_NORM;
        j->msg = "norm more than 1";
    } else {
        j->state = kDone;
    }
    j->t_done = p->now();
    qjx_ctx_destroy(c);
    j->ctx = nullptr;
    p->used -= j->footprint;
    p->in_table -= j->e.num_trajectories;
    return QJX_OK;
}

static int sync_table(qjx_pool *p)
{
    std::vector<qjx::Ens> tab;
    for (int id : p->running) tab.push_back(p->jobs[id]->ctx->desc);
    p->table_dirty = false;
    return qjx::runner_set_table(p->runner, tab.data(), (int)tab.size());
}

// One scheduling cycle: launch a chunk of iterations over the running jobs,
// admit queued jobs while it runs, then retire the jobs that finished.
extern "C" int qjx_pool_step(qjx_pool *p, int *running, int *queued)
{
    if (!p) return pool_fail(QJX_ERR_ARG, "null pool");
    int rc = QJX_OK;
    if (p->table_dirty && (rc = sync_table(p))) return rc;
    const bool busy = !p->running.empty();
    // the launch shape is tuned once, after the first jobs' bisections have spread
    if (busy && p->cycles >= 4 && (rc = qjx::runner_autotune(p->runner))) return rc;
    // no graph capture while the builder thread may call the runtime: its calls can
    // invalidate a capture in progress.  An idle builder only waits on `cv`, and no
    // build starts before this chunk is launched.
    {
        std::lock_guard<std::mutex> lk(p->mu);
        p->runner->defer_capture = p->building > 0;
    }
    if (busy && (rc = qjx::runner_chunk(p->runner, 16))) return rc;
    if (busy) p->cycles++;
    // built jobs join the table at the next cycle; admission reserves budget and
    // hands the build to the builder thread (it overlaps the chunk just launched)
    {
        std::lock_guard<std::mutex> lk(p->mu);
        while (!p->built.empty()) {
            Job *j = p->jobs[p->built.front()];
            p->built.pop_front();
            p->building--;
            if (j->err) {
                j->state = kFailed;
                p->used -= j->footprint;
                p->in_table -= j->e.num_trajectories;
                continue;
            }
            j->state = kRunning;
            p->running.push_back(j->id);
            p->table_dirty = true;
        }
        while (!p->queue.empty() && (int)p->running.size() + p->building < qjx::kMaxEns) {
            Job *j = p->jobs[p->queue.front()];
            if (p->used + j->footprint > p->budget) break;
            if (p->in_table + j->e.num_trajectories > p->capacity) break;
            p->queue.pop_front();
            p->used += j->footprint;
            p->in_table += j->e.num_trajectories;
            j->t_admit = p->now();
            p->to_build.push_back(j->id);
            p->building++;
        }
    }
    p->cv.notify_one();
    if (busy) {
        if (hipStreamSynchronize(p->stream) != hipSuccess) return pool_fail(QJX_ERR_HIP, "pool iteration failed");
        // a job is finished when its lists for the next iteration are empty
        std::vector<int> still;
        for (int id : p->running) {
            Job *j = p->jobs[id];
            int cnt[qjx::kMaxGroups];
            if (hipMemcpy(cnt, j->ctx->cnts[0], sizeof cnt, hipMemcpyDeviceToHost) != hipSuccess)
                return pool_fail(QJX_ERR_HIP, "pool status read failed");
            long long left = 0;
            for (int g = 0; g < j->ctx->G; g++) left += cnt[g];
            // jobs admitted during this cycle were seeded, so their lists are not empty
            if (left == 0) {
                rc = retire(p, j);
                if (rc) return rc;
                p->table_dirty = true;
            } else {
                still.push_back(id);
            }
        }
        p->running.swap(still);
    }
    if (running) *running = (int)p->running.size();
    if (queued) {
        std::lock_guard<std::mutex> lk(p->mu);
        *queued = (int)p->queue.size() + p->building;
    }
    return QJX_OK;
}

extern "C" int qjx_pool_result(qjx_pool *p, int job_id, qjx_job_result *out, double *mean, double *energy)
{
    if (!p || !out || job_id < 0 || job_id >= (int)p->jobs.size()) return pool_fail(QJX_ERR_ARG, "bad job id");
    const Job *j = p->jobs[job_id];
    std::memset(out, 0, sizeof *out);
    out->status = j->state == kFailed ? j->err : j->state;
    out->t_submit = j->t_submit;
    out->t_admit = j->t_admit;
    out->t_done = j->t_done;
    out->num_trajectories = j->e.num_trajectories;
    out->n_periods = j->n_periods;
    out->dim = j->m.N + 1;
    out->matvecs = j->matvecs;
    out->jumps = j->jumps;
    if (j->state == kDone) {
        const int M = j->e.num_trajectories;
        double sm = 0, sm2 = 0, se = 0, se2 = 0, sn = 0;
        for (int i = 0; i < M; i++) {
            sm += j->mean[i];
            sm2 += j->mean[i] * j->mean[i];
            se += j->energy[i];
            se2 += j->energy[i] * j->energy[i];
            sn += j->norm[i];
        }
        out->mean_avg = sm / M;
        out->mean_std = std::sqrt(std::max(0.0, sm2 / M - out->mean_avg * out->mean_avg));
        out->energy_avg = se / M;
        out->energy_std = std::sqrt(std::max(0.0, se2 / M - out->energy_avg * out->energy_avg));
        out->norm_avg = sn / M;
        if (mean) std::memcpy(mean, j->mean.data(), M * sizeof(double));
        if (energy) std::memcpy(energy, j->energy.data(), M * sizeof(double));
    }
    return QJX_OK;
}
