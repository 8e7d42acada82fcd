// qjx_lpn.cpp -- the LPN task (os_d task_id 7, LpnMultExperimentBehaviour) over a
// trajectory context: Lyapunov-type growth rates of observable distances between
// paired trajectories.
//
// Reference (source/cpp/QJX/QJX/): LpnMultNewDelBehaviour::init_data
// (newdel_exp.cpp:29-48), init_streams_var (newdel_exp.cpp:313-330), init_obs_lpn
// (newdel_exp.cpp:528-600), LpnMultExperimentBehaviour::trans_process / obser_process
// (experiment.cpp:137-286), copy_trajectory_lpn / copy_stream_lpn / var_trajectory_lpn
// (experiment.cpp:2221-2393), lambda_lpn / lambda_lpn_now (experiment.cpp:3065-3134),
// calc_chars_lpn(_start) / calc_delta_s / calc_delta_f (core.cpp:727-915),
// init_random_obs / get_random_obs (newdel.cpp:3652-3722, experiment.cpp:2023-2078).
//
// Trajectory tr < M/2 is an original; tr + M/2 is its partner: after the transient it
// takes the original's state, RNG stream and eta, plus a perturbation of chosen
// observable distance.  Every period the GPU advances all M trajectories and then,
// on the device-resident states, evaluates every trajectory's observables, the pair
// distances, and renormalises the partners (a fresh perturbation of the original and
// the log growth): no state leaves the device, and the host reads scalars only.
//
// The device code (qjx_obs.hip) performs the CPU restatement's operations in its
// order (the reference's loops; the banded energy adds the same nonzero terms as its
// dense loop; 10^y and ln x by the same series on both sides), so with bit-identical
// propagation (the RK path) the LPN results are bit-identical too.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/qjx.h"
#include "qjx_ctx.h"
#include "qjx_obs.h"

namespace {

int lpn_fail(int code, const std::string &msg)
{
    qjx::set_last_error(msg);
    return code;
}

}  // namespace

struct qjx_lpn {
    qjx_ctx *c = nullptr;
    qjx_lpn_desc p{};
    int M = 0, half = 0, d = 0, period = 0;
    qjx::LpnArgs a{};
    std::vector<void *> allocs;
    double2 *R = nullptr;
    std::vector<std::vector<double>> lambdas, deltas_s, deltas_f;   // save_lambdas records

    ~qjx_lpn()
    {
        if (c && c->stream) (void)hipStreamSynchronize(c->stream);
        for (void *q : allocs) (void)hipFree(q);
        if (R) (void)hipFree(R);
    }

    template <typename T>
    int alloc(T **out, size_t n)
    {
        void *q = nullptr;
        if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess ||
            hipMemsetAsync(q, 0, std::max<size_t>(n, 1) * sizeof(T), c->stream) != hipSuccess)
            return lpn_fail(QJX_ERR_NOMEM, "LPN buffers do not fit in device memory");
        allocs.push_back(q);
        *out = static_cast<T *>(q);
        return QJX_OK;
    }

    // the context's current state buffers (an ensemble_init reallocates them)
    void refresh()
    {
        a.slot0 = c->slot[0];
        a.slot1 = c->slot[1];
        a.st = c->st;
    }

    int sync()
    {
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess)
            return lpn_fail(QJX_ERR_HIP, "LPN launch failed");
        return QJX_OK;
    }

    // device -> host in the context's stream order
    template <typename T>
    int get(const T *src, T *dst, size_t n)
    {
        if (!dst) return QJX_OK;
        if (hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            return lpn_fail(QJX_ERR_HIP, "LPN readback failed");
        return QJX_OK;
    }
};

extern "C" int qjx_lpn_create(qjx_ctx *c, const qjx_lpn_desc *desc, qjx_lpn **out)
{
    if (!c || !desc || !out) return lpn_fail(QJX_ERR_ARG, "null argument");
    *out = nullptr;
    if (!c->have_ens) return lpn_fail(QJX_ERR_ARG, "ensemble not initialised");
    if (c->M < 2 || (c->M % 2)) return lpn_fail(QJX_ERR_ARG, "LPN needs an even number of trajectories");
    if (desc->lpn_type < -2 || desc->lpn_type >= desc->num_random_obs)
        return lpn_fail(QJX_ERR_ARG, "wrong lpn_type (-1, -2 or a random observable index)");
    if (!(desc->lpn_eps_error > 0.0)) return lpn_fail(QJX_ERR_ARG, "lpn_eps_error must be > 0");
    if (desc->num_random_obs < 0) return lpn_fail(QJX_ERR_ARG, "num_random_obs must be >= 0");
    qjx_lpn *L = new qjx_lpn;
    L->c = c;
    L->p = *desc;
    L->M = c->M;
    L->half = c->M / 2;
    L->d = c->d;
    const int d = c->d, M = c->M, half = c->M / 2, no = desc->num_random_obs;
    auto bail = [&](int rc) {
        delete L;
        return rc;
    };
    qjx::LpnArgs &a = L->a;
    a.M = M;
    a.half = half;
    a.d = d;
    a.dp = c->dp;
    a.L = c->L;
    a.prec = c->prec;
    a.num = no;
    a.E0 = c->model.prm_E;
    a.T = c->T;
    a.max_energy = 0.0;
    for (int i = 0; i < d; i++) a.max_energy = std::max(a.max_energy, std::fabs(c->Hdiag[i]));   // init_obs_lpn
    a.lpn_type = desc->lpn_type;
    a.eps_deep = desc->lpn_eps_deep;
    a.lambda_per_periods = desc->lambda_per_periods;
    a.save = desc->save_lambdas > 0;
    a.eps_error = desc->lpn_eps_error;
    a.eps_high = desc->lpn_eps_high;
    a.eps_low = desc->lpn_eps_low;
    a.delta_s_target = desc->lpn_delta_s;
    a.df_high = desc->lpn_delta_f_high;
    a.df_low = desc->lpn_delta_f_low;
    int rc = qjx::build_random_obs(d, no, desc->random_obs_seed, desc->random_obs_mns, desc->random_obs_type, &L->R,
                                   c->stream);
    if (rc) return bail(rc);
    a.R = L->R;
    double *Hd = nullptr, *Ho = nullptr, *Hv = nullptr;
    uint64_t *vx = nullptr;
    if ((rc = L->alloc(&Hd, d)) || (rc = L->alloc(&Ho, d)) || (rc = L->alloc(&Hv, d)) ||
        (rc = L->alloc(&a.norm, M)) || (rc = L->alloc(&a.mean, M)) || (rc = L->alloc(&a.disp, M)) ||
        (rc = L->alloc(&a.m2, M)) || (rc = L->alloc(&a.energy, M)) || (rc = L->alloc(&a.mean_start, M)) ||
        (rc = L->alloc(&a.mean_lpn, M)) || (rc = L->alloc(&a.energy_lpn, M)) || (rc = L->alloc(&a.lambda, M)) ||
        (rc = L->alloc(&a.lambda_now, M)) || (rc = L->alloc(&a.delta_s, M)) ||
        (rc = L->alloc(&a.num_renorms, M)) || (rc = L->alloc(&a.obs, (size_t)M * no)) ||
        (rc = L->alloc(&a.obs_lpn, (size_t)M * no)) || (rc = L->alloc(&a.adr, (size_t)M * d)) ||
        (rc = L->alloc(&vx, M)) || (rc = L->alloc(&a.work, (size_t)half * 2 * d)) ||
        (rc = L->alloc(&a.rec_flag, M)) || (rc = L->alloc(&a.rec_ratio, M)) || (rc = L->alloc(&a.rec_ds, M)) ||
        (rc = L->alloc(&a.rec_df, M)))
        return bail(rc);
    // H, H_drv on their band as the energy loop reads them
    std::vector<double> hoff(d, 0.0);
    for (int k = 0; k + 1 < d; k++) hoff[k] = c->Hoff[k];
    // streams_var: MCG31(777) leap-frog(tr, M) (init_streams_var, newdel_exp.cpp:313-330)
    std::vector<uint64_t> x0(M);
    uint64_t stride = 1;
    for (int tr = 0; tr < M; tr++) qjx::mcg31_stream(777, tr, M, &x0[tr], &stride);
    a.var_stride = stride;
    // uploads in the context's stream order, after the buffers' clearing (the host
    // vectors live until the sync below)
    if (hipMemcpyAsync(Hd, c->Hdiag.data(), d * sizeof(double), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(Ho, hoff.data(), d * sizeof(double), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(Hv, c->Hdrv.data(), d * sizeof(double), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(vx, x0.data(), M * sizeof(uint64_t), hipMemcpyHostToDevice, c->stream) != hipSuccess)
        return bail(lpn_fail(QJX_ERR_HIP, "LPN upload failed"));
    a.Hdiag = Hd;
    a.Hoff = Ho;
    a.Hdrv = Hv;
    a.var_x = vx;
    L->lambdas.assign(M, {});
    L->deltas_s.assign(M, {});
    L->deltas_f.assign(M, {});
    if ((rc = L->sync())) return bail(rc);
    *out = L;
    return QJX_OK;
}

extern "C" void qjx_lpn_destroy(qjx_lpn *L) { delete L; }

// LpnMultExperimentBehaviour::trans_process: the transient of the originals (the
// partners' transient runs too and is overwritten), then each partner copies its
// original and is perturbed; every time restarts at 0.
extern "C" int qjx_lpn_trans(qjx_lpn *L, int num_tp_periods)
{
    if (!L) return lpn_fail(QJX_ERR_ARG, "null LPN handle");
    int rc = qjx_trans(L->c, num_tp_periods);
    if (rc) return rc;
    L->refresh();
    hipStream_t s = L->c->stream;
    qjx::launch_lpn_chars_std(L->a, 0, L->half, 1, s);
    qjx::launch_lpn_start(L->a, 0, L->half, s);
    qjx::launch_lpn_pair_start(L->a, s);
    qjx::launch_lpn_chars_std(L->a, L->half, L->half, 1, s);
    qjx::launch_lpn_start(L->a, L->half, L->half, s);
    if ((rc = L->sync())) return rc;
    L->period = 0;
    return QJX_OK;
}

// n observation periods (obser_process's period loop): all trajectories advance one
// period, then every trajectory's observables, the pair distances and, past a
// threshold (or every period with lambda_per_periods), the renormalisations run on
// the device.
extern "C" int qjx_lpn_run_periods(qjx_lpn *L, int n)
{
    if (!L) return lpn_fail(QJX_ERR_ARG, "null LPN handle");
    if (n < 0) return lpn_fail(QJX_ERR_ARG, "negative period count");
    const int M = L->M, half = L->half;
    std::vector<int> flag;
    std::vector<double> ratio, ds, df;
    for (int k = 0; k < n; k++) {
        int rc = qjx_run_periods(L->c, 1, 1);
        if (rc) return rc;
        L->refresh();
        hipStream_t s = L->c->stream;
        qjx::launch_lpn_chars_std(L->a, 0, M, 0, s);
        qjx::launch_lpn_chars_lpn(L->a, 0, half, s);
        qjx::launch_lpn_pairs(L->a, double(L->period + 1) * L->c->T, s);
        if ((rc = L->sync())) return rc;
        if (L->a.save) {
            flag.resize(M);
            ratio.resize(M);
            ds.resize(M);
            df.resize(M);
            if ((rc = L->get<int>(L->a.rec_flag, flag.data(), M)) ||
                (rc = L->get<double>(L->a.rec_ratio, ratio.data(), M)) ||
                (rc = L->get<double>(L->a.rec_ds, ds.data(), M)) || (rc = L->get<double>(L->a.rec_df, df.data(), M)))
                return rc;
            for (int tr = half; tr < M; tr++)
                if (flag[tr]) {
                    L->lambdas[tr].push_back(ratio[tr]);
                    L->deltas_s[tr].push_back(ds[tr]);
                    L->deltas_f[tr].push_back(df[tr]);
                }
        }
        L->period++;
    }
    return QJX_OK;
}

extern "C" int qjx_lpn_get(qjx_lpn *L, double *lambda_now, double *lambda, int *num_renorms, double *mean_lpn,
                           double *energy_lpn, double *delta_s)
{
    if (!L) return lpn_fail(QJX_ERR_ARG, "null LPN handle");
    const size_t M = L->M;
    int rc;
    if ((rc = L->get(L->a.lambda_now, lambda_now, M)) || (rc = L->get(L->a.lambda, lambda, M)) ||
        (rc = L->get(L->a.mean_lpn, mean_lpn, M)) || (rc = L->get(L->a.energy_lpn, energy_lpn, M)) ||
        (rc = L->get(L->a.delta_s, delta_s, M)))
        return rc;
    return L->get<int>(L->a.num_renorms, num_renorms, M);
}

extern "C" int qjx_lpn_random_obs(qjx_lpn *L, double *obs, double *obs_lpn)
{
    if (!L) return lpn_fail(QJX_ERR_ARG, "null LPN handle");
    const size_t n = (size_t)L->M * L->a.num * 2;
    if (n == 0) return QJX_OK;
    int rc;
    if ((rc = L->get(reinterpret_cast<const double *>(L->a.obs), obs, n)) ||
        (rc = L->get(reinterpret_cast<const double *>(L->a.obs_lpn), obs_lpn, n)))
        return rc;
    return QJX_OK;
}

extern "C" int qjx_lpn_lambdas(qjx_lpn *L, int tr, double *ratio, double *ds, double *df, int cap)
{
    if (!L) return lpn_fail(QJX_ERR_ARG, "null LPN handle");
    if (tr < 0 || tr >= L->M) return lpn_fail(QJX_ERR_ARG, "trajectory out of range");
    const int n = (int)L->lambdas[tr].size();
    for (int i = 0; i < n && i < cap; i++) {
        if (ratio) ratio[i] = L->lambdas[tr][i];
        if (ds) ds[i] = L->deltas_s[tr][i];
        if (df) df[i] = L->deltas_f[tr][i];
    }
    return n;
}

extern "C" int qjx_lpn_random_obs_matrix(qjx_lpn *L, int obs, double *R)
{
    if (!L || !R) return lpn_fail(QJX_ERR_ARG, "null argument");
    if (obs < 0 || obs >= L->a.num) return lpn_fail(QJX_ERR_ARG, "random observable out of range");
    const size_t dd = (size_t)L->d * L->d;
    return L->get(reinterpret_cast<const double *>(L->R + obs * dd), R, 2 * dd);
}

// The last period's calc_chars_std values (before any renormalisation of that
// period), the current abs_diag_rho (M x d) and its average over trajectories
// [begin, end) as dump_adr_avg(_mult) forms it (dump.cpp:26-98).  Any pointer may
// be NULL.
extern "C" int qjx_lpn_chars(qjx_lpn *L, double *norm, double *mean, double *dispersion, double *m2,
                             double *energy, double *adr, int begin, int end, double *adr_avg)
{
    if (!L) return lpn_fail(QJX_ERR_ARG, "null LPN handle");
    if (adr_avg && (begin < 0 || end > L->M || end <= begin)) return lpn_fail(QJX_ERR_ARG, "bad trajectory range");
    const int M = L->M, d = L->d;
    int rc;
    if ((rc = L->get(L->a.norm, norm, M)) || (rc = L->get(L->a.mean, mean, M)) ||
        (rc = L->get(L->a.disp, dispersion, M)) || (rc = L->get(L->a.m2, m2, M)) ||
        (rc = L->get(L->a.energy, energy, M)))
        return rc;
    if (adr || adr_avg) {
        std::vector<double> h((size_t)M * d);
        if ((rc = L->get(L->a.adr, h.data(), h.size()))) return rc;
        if (adr) std::memcpy(adr, h.data(), sizeof(double) * h.size());
        if (adr_avg) {
            const double n = double(end - begin);
            for (int k = 0; k < d; k++) adr_avg[k] = 0.0;
            for (int tr = begin; tr < end; tr++)
                for (int k = 0; k < d; k++) adr_avg[k] += h[(size_t)tr * d + k] / n;
        }
    }
    return QJX_OK;
}
