// qjx_pool.cpp -- continuous batching of ensemble jobs on one GPU.
//
// A job is one dimer parameter point with its trajectory ensemble and a number of
// periods (the reference runs one such ensemble per process, one OpenMP loop per
// period, experiment.cpp:448-542; parameter sweeps are separate jobs,
// submit/qjx/submit_dimer.py).  Here every running job is one entry of a shared
// ensemble table: each propagation iteration is one set of launches over the
// trajectories of all running jobs of a lane, whatever their Hilbert dimension.
//
// Scheduling:
//   * Admission takes queued jobs in order of expected device work (shortest first)
//     and skips a job that does not fit rather than blocking the ones behind it.
//     A job is admitted while its HBM footprint fits the budget and its
//     trajectories fit the lane's capacity.
//   * Two lanes, each its own table and runner: short jobs (expected device work
//     below a threshold) and long ones.  The short lane takes in a bounded number
//     of trajectories (QJX_POOL_SHORT_TRAJ, default 1e5) so that its iterations
//     stay short, and gets 3/4 of the device time when both lanes have work
//     (QJX_POOL_SHORT_SHARE).  An iteration's time is set by the largest
//     jobs in its table, and a trajectory needs one iteration per matrix-vector
//     step of its bisection chain, so a short job sharing iterations with hour-long
//     jobs would wait on their GEMMs at every step.  In its own lane it iterates at
//     its own pace.  Long jobs may not take the budget share reserved for short ones.
//   * Models are built by several builder threads, each on the job's own stream,
//     while the lanes iterate.
// A trajectory's arithmetic does not depend on which other jobs share the
// launches or in which lane it runs, so a job's results are the same alone and in
// the mix.
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
enum Lane { kShort = 0, kLong = 1, kLanes = 2 };

struct Job {
    int id = 0;
    qjx_model_desc m{};
    qjx_ensemble_desc e{};
    int n_periods = 0;
    int state = kQueued;
    int lane = kShort;
    int err = 0;
    std::string msg;
    qjx_ctx *ctx = nullptr;
    size_t footprint = 0;
    double work = 0.0;                // expected device seconds (admission order, lane)
    double t_submit = 0, t_admit = 0, t_done = 0;
    std::vector<double> norm, mean, m2, energy;
    long long matvecs = 0, jumps = 0;
};

int pool_fail(int code, const std::string &msg)
{
    qjx::set_last_error(msg);
    return code;
}

// Expected device seconds of a job: M P (S + 2 D J) propagator applications of
// 8 d^2 flops, J the expected jumps per period (half the largest rate
// gamma (N + 2) times T), at a dense-equivalent rate of 3e14 flop/s.  Only its
// order matters: it ranks jobs for admission and picks their lane.
double expected_work(const qjx_model_desc &m, int M, int periods)
{
    const double T = 2.0 * 3.141592653589793 / m.drv_freq;
    const double J = 0.5 * m.diss_gamma * (m.N + 2) * T;
    const int S = m.steps_per_period ? m.steps_per_period : 4;
    const double d = m.N + 1;
    return (double)M * periods * (S + 2.0 * m.ex_deep * J) * 8.0 * d * d / 3e14;
}

double env_double(const char *name, double dflt)
{
    const char *v = std::getenv(name);
    return v ? std::atof(v) : dflt;
}

}  // namespace

struct qjx_pool {
    int prec = 0;
    int gemm_kind = qjx::kGemmMfma;
    size_t budget = 0, used = 0, used_long = 0;
    int capacity = 0;
    hipStream_t stream = nullptr;
    struct LaneState {
        qjx::Runner *runner = nullptr;
        std::vector<int> running;     // ids in table order
        bool dirty = false;
        long long cycles = 0;         // cycles that launched iterations in this lane
        long long in_table = 0;       // trajectories admitted to this lane
        int pending = 0;              // admitted jobs not yet running (building or built)
        int iters = 16;               // iterations of the lane's next chunk
        hipEvent_t ev0 = nullptr, ev1 = nullptr;
    } lane[kLanes];
    double short_s = 0.5;             // expected seconds below which a job is short
    double reserve = 0.25;            // budget share long jobs may not take
    long long short_cap = 100000;     // trajectories the short lane takes in (a larger job: alone)
    double short_share = 0.75;        // device time of the short lane when both lanes have work
    std::vector<Job *> jobs;          // by id
    std::vector<int> queue;           // waiting for admission
    std::vector<std::thread> builders;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<int> to_build;         // admitted (budget reserved), not yet built
    std::deque<int> built;            // built and seeded, waiting to join a lane
    int building = 0;                 // jobs handed to the builders and not yet in `built`
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

static void destroy_lanes(qjx_pool *p)
{
    for (auto &L : p->lane) {
        qjx::runner_destroy(L.runner);
        L.runner = nullptr;
        if (L.ev0) (void)hipEventDestroy(L.ev0);
        if (L.ev1) (void)hipEventDestroy(L.ev1);
        L.ev0 = L.ev1 = nullptr;
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
    p->short_s = env_double("QJX_POOL_SHORT_S", 0.5);
    p->reserve = std::min(0.9, std::max(0.0, env_double("QJX_POOL_RESERVE", 0.25)));
    p->short_cap = std::max(1LL, (long long)env_double("QJX_POOL_SHORT_TRAJ", 1e5));
    p->short_share = std::min(0.95, std::max(0.05, env_double("QJX_POOL_SHORT_SHARE", 0.75)));
    size_t freeb = 0, totalb = 0;
    if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) {
        delete p;
        return pool_fail(QJX_ERR_HIP, "no HIP device available");
    }
    p->budget = hbm_budget_gib > 0 ? (size_t)(hbm_budget_gib * (1ULL << 30)) : (size_t)(0.8 * (double)freeb);
    // Non-blocking: the builder threads' null-stream calls (hipMemcpy, hipMemset,
    // hipMalloc) then neither wait for the running chunk nor invalidate a graph
    // capture in progress on this stream.  Every hand-over between them is ordered
    // by the host instead: a build ends with hipStreamSynchronize on the job's own
    // stream, and results are read after synchronising this one.
    if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
        delete p;
        return pool_fail(QJX_ERR_HIP, "hipStreamCreate failed");
    }
    bool graphs = true;
    if (const char *env = std::getenv("QJX_GRAPH"))
        if (std::string(env) == "0") graphs = false;
    for (auto &L : p->lane) {
        int rc = qjx::runner_create(precision, p->gemm_kind, p->stream, max_trajectories, &L.runner);
        if (!rc && (hipEventCreate(&L.ev0) != hipSuccess || hipEventCreate(&L.ev1) != hipSuccess))
            rc = pool_fail(QJX_ERR_HIP, "hipEventCreate failed");
        if (rc) {
            destroy_lanes(p);
            (void)hipStreamDestroy(p->stream);
            delete p;
            return rc;
        }
        L.runner->graphs = graphs;
    }
    (void)hipGetDevice(&p->device);
    int nb = (int)env_double("QJX_POOL_BUILDERS", 4.0);
    nb = std::max(1, std::min(16, nb));
    for (int i = 0; i < nb; i++) p->builders.emplace_back(builder_main, p);
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
    for (auto &t : p->builders)
        if (t.joinable()) t.join();
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    for (Job *j : p->jobs) {
        if (j->ctx) qjx_ctx_destroy(j->ctx);
        delete j;
    }
    destroy_lanes(p);
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
    j->m = *m;
    j->m.precision = p->prec;
    j->e = *e;
    j->e.jump = 0;          // serving jobs return observables, not jump logs
    j->n_periods = n_periods;
    j->footprint = qjx::model_footprint(&j->m, p->gemm_kind) +
                   qjx::ensemble_footprint(&j->m, e->num_trajectories, p->gemm_kind);
    j->work = expected_work(j->m, e->num_trajectories, n_periods);
    j->lane = j->work <= p->short_s ? kShort : kLong;
    const double room = j->lane == kShort ? 1.0 : 1.0 - p->reserve;
    if ((double)j->footprint > room * (double)p->budget) {
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

// Release a job's reservation (budget, lane capacity); the caller holds p->mu.
static void release(qjx_pool *p, Job *j)
{
    p->used -= j->footprint;
    if (j->lane == kLong) p->used_long -= j->footprint;
    p->lane[j->lane].in_table -= j->e.num_trajectories;
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
    const bool norm_err = hipMemcpy(h, c->err, sizeof h, hipMemcpyDeviceToHost) == hipSuccess && h[0];
    qjx_ctx_destroy(c);
    std::lock_guard<std::mutex> lk(p->mu);
    if (norm_err) {
        j->state = kFailed;
        j->err = QJX_ERR_NORM;
        j->msg = "norm more than 1";
    } else {
        j->state = kDone;
    }
    j->t_done = p->now();
    j->ctx = nullptr;
    release(p, j);
    return QJX_OK;
}

static int sync_table(qjx_pool *p, int l)
{
    auto &L = p->lane[l];
    std::vector<qjx::Ens> tab;
    for (int id : L.running) tab.push_back(p->jobs[id]->ctx->desc);
    L.dirty = false;
    return qjx::runner_set_table(L.runner, tab.data(), (int)tab.size());
}

// Admission (caller holds p->mu): queued jobs in order of expected work; a job
// that does not fit is skipped, not waited for.
static void admit_queued(qjx_pool *p)
{
    if (p->queue.empty()) return;
    std::stable_sort(p->queue.begin(), p->queue.end(),
                     [&](int a, int b) { return p->jobs[a]->work < p->jobs[b]->work; });
    int tables[kLanes];
    for (int l = 0; l < kLanes; l++) tables[l] = (int)p->lane[l].running.size() + p->lane[l].pending;
    const size_t long_cap = (size_t)((1.0 - p->reserve) * (double)p->budget);
    std::vector<int> still;
    for (size_t k = 0; k < p->queue.size(); k++) {
        Job *j = p->jobs[p->queue[k]];
        const int l = j->lane;
        // the short lane's trajectories are capped: its iteration time, which every
        // short job pays once per step of its bisection chain, grows with them (a job
        // above the cap runs when the lane is empty)
        const long long lane_cap = l == kShort ? std::min<long long>(p->capacity, p->short_cap) : p->capacity;
        const bool room = p->lane[l].in_table + j->e.num_trajectories <= lane_cap ||
                          (p->lane[l].in_table == 0 && tables[l] == 0);
        const bool fits = p->used + j->footprint <= p->budget &&
                          (l == kShort || p->used_long + j->footprint <= long_cap) &&
                          room && p->lane[l].in_table + j->e.num_trajectories <= p->capacity &&
                          tables[l] < qjx::kMaxEns;
        if (!fits) {
            still.push_back(j->id);
            continue;
        }
        p->used += j->footprint;
        if (l == kLong) p->used_long += j->footprint;
        p->lane[l].in_table += j->e.num_trajectories;
        p->lane[l].pending++;
        tables[l]++;
        j->t_admit = p->now();
        p->to_build.push_back(j->id);
        p->building++;
    }
    p->queue.swap(still);
}

// One scheduling cycle: launch a chunk of iterations in each busy lane, admit
// queued jobs while it runs, then retire the jobs that finished.
extern "C" int qjx_pool_step(qjx_pool *p, int *running, int *queued)
{
    if (!p) return pool_fail(QJX_ERR_ARG, "null pool");
    int rc = QJX_OK;
    bool busy[kLanes];
    bool deferred;
    {
        std::lock_guard<std::mutex> lk(p->mu);
        // no graph capture while builder threads may call the runtime: their calls
        // can invalidate a capture in progress.  Idle builders only wait on `cv`.
        deferred = p->building > 0;
    }
    for (int l = 0; l < kLanes; l++) {
        auto &L = p->lane[l];
        if (L.dirty && (rc = sync_table(p, l))) return rc;
        busy[l] = !L.running.empty();
        // the launch shape is tuned once, after the first jobs' bisections have spread
        if (busy[l] && L.cycles >= 4 && (rc = qjx::runner_autotune(L.runner))) return rc;
        L.runner->defer_capture = deferred;
    }
    for (int l = 0; l < kLanes; l++) {
        if (!busy[l]) continue;
        auto &L = p->lane[l];
        if (hipEventRecord(L.ev0, p->stream) != hipSuccess) return pool_fail(QJX_ERR_HIP, "event record failed");
        if ((rc = qjx::runner_chunk(L.runner, L.iters))) return rc;
        if (hipEventRecord(L.ev1, p->stream) != hipSuccess) return pool_fail(QJX_ERR_HIP, "event record failed");
        L.cycles++;
    }
    // built jobs join their lane at the next cycle; admission reserves budget and
    // hands the builds to the builder threads (they overlap the chunks just launched)
    {
        std::lock_guard<std::mutex> lk(p->mu);
        while (!p->built.empty()) {
            Job *j = p->jobs[p->built.front()];
            p->built.pop_front();
            p->building--;
            p->lane[j->lane].pending--;
            if (j->err) {
                j->state = kFailed;
                release(p, j);
                continue;
            }
            j->state = kRunning;
            p->lane[j->lane].running.push_back(j->id);
            p->lane[j->lane].dirty = true;
        }
        admit_queued(p);
    }
    p->cv.notify_all();
    if (busy[kShort] || busy[kLong]) {
        if (hipStreamSynchronize(p->stream) != hipSuccess) return pool_fail(QJX_ERR_HIP, "pool iteration failed");
        // When both lanes have work the short lane gets short_share of the device time
        // (default 3/4: its jobs can finish within minutes, the long lane's take hours):
        // the long lane's chunk stays short enough to keep cycles near 0.2 s, and the
        // short lane's chunk is sized to last share / (1 - share) times as long.
        float ms[kLanes] = {0.f, 0.f};
        for (int l = 0; l < kLanes; l++)
            if (busy[l] && hipEventElapsedTime(&ms[l], p->lane[l].ev0, p->lane[l].ev1) != hipSuccess)
                return pool_fail(QJX_ERR_HIP, "event timing failed");
        auto even = [](double x) { return std::max(2, (int)(x / 2.0) * 2); };
        auto &S = p->lane[kShort], &Lg = p->lane[kLong];
        if (busy[kLong] && ms[kLong] > 0.f) {
            const double per = ms[kLong] / Lg.iters;
            Lg.iters = std::min(16, even(200.0 / std::max(per, 1e-3)));
        }
        if (busy[kShort] && ms[kShort] > 0.f) {
            const double per = ms[kShort] / S.iters;
            const double ratio = p->short_share / (1.0 - p->short_share);
            const double target = busy[kLong] ? std::max((double)ms[kLong] * ratio, 2.0) : 16.0 * per;
            S.iters = std::min(1024, std::max(16, even(target / std::max(per, 1e-3))));
        }
        // a job is finished when its lists for the next iteration are empty
        for (int l = 0; l < kLanes; l++) {
            if (!busy[l]) continue;
            auto &L = p->lane[l];
            std::vector<int> still;
            for (int id : L.running) {
                Job *j = p->jobs[id];
                int cnt[qjx::kMaxGroups];
                if (hipMemcpy(cnt, j->ctx->cnts[0], sizeof cnt, hipMemcpyDeviceToHost) != hipSuccess)
                    return pool_fail(QJX_ERR_HIP, "pool status read failed");
                long long left = 0;
                for (int g = 0; g < j->ctx->G; g++) left += cnt[g];
                // jobs that joined at this cycle were seeded, so their lists are not empty
                if (left == 0) {
                    if ((rc = retire(p, j))) return rc;
                    L.dirty = true;
                } else {
                    still.push_back(id);
                }
            }
            L.running.swap(still);
        }
    }
    std::lock_guard<std::mutex> lk(p->mu);
    if (running) *running = (int)(p->lane[kShort].running.size() + p->lane[kLong].running.size());
    if (queued) *queued = (int)p->queue.size() + p->building;
    return QJX_OK;
}

extern "C" int qjx_pool_result(qjx_pool *p, int job_id, qjx_job_result *out, double *mean, double *energy)
{
    if (!p || !out) return pool_fail(QJX_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(p->mu);
    if (job_id < 0 || job_id >= (int)p->jobs.size()) return pool_fail(QJX_ERR_ARG, "bad job id");
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

extern "C" int qjx_pool_get_usage(qjx_pool *p, qjx_pool_usage *out)
{
    if (!p || !out) return pool_fail(QJX_ERR_ARG, "null argument");
    std::memset(out, 0, sizeof *out);
    std::lock_guard<std::mutex> lk(p->mu);
    out->used_bytes = (int64_t)p->used;
    out->budget_bytes = (int64_t)p->budget;
    out->trajectories = p->lane[kShort].in_table + p->lane[kLong].in_table;
    out->capacity = p->capacity;
    out->queued = (int)p->queue.size();
    out->building = p->building;
    out->running = (int)(p->lane[kShort].running.size() + p->lane[kLong].running.size());
    for (const Job *j : p->jobs) {
        if (j->state == kDone) out->done++;
        if (j->state == kFailed) out->failed++;
    }
    return QJX_OK;
}
