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
// observable distance.  Every period the GPU advances all M trajectories; the host
// then reads the states back and does the per-pair bookkeeping (observable distance,
// renormalisation by a fresh perturbation of the original, log growth).  That
// bookkeeping is O(M d) per period (O(M d^2) per random observable), the propagation
// O(M d^2 steps), so the split keeps the device busy with the part that scales.
//
// The observables used for distances are computed here in the same arithmetic as
// the CPU restatement (the reference's loops; the banded energy adds the same nonzero
// terms as its dense loop), so with bit-identical propagation (the RK path) the LPN
// results are bit-identical too.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/qjx.h"
#include "qjx_ctx.h"

#pragma clang fp contract(off)

namespace {

int lpn_fail(int code, const std::string &msg)
{
    qjx::set_last_error(msg);
    return code;
}

const double kPi = 3.1415926535897932384626433832795;   // reference config.h PI

// MKL VSL MCG31m1: x <- a x mod (2^31 - 1), a = 1132489760; a uniform on [lo, hi)
// is fma(double(x), (hi - lo) (1 / m), lo), after which the stream advances.
const uint64_t kMcg31A = 1132489760ULL, kMcg31M = 2147483647ULL;

struct Mcg31 {
    uint64_t x = 1, stride = 1;
    double next(double lo, double hi)
    {
        const double r = std::fma(double(x), (hi - lo) * (1.0 / double(kMcg31M)), lo);
        x = x * stride % kMcg31M;
        return r;
    }
};

uint64_t mcg31_pow(uint64_t a, uint64_t e)
{
    uint64_t r = 1;
    a %= kMcg31M;
    while (e) {
        if (e & 1ULL) r = r * a % kMcg31M;
        a = a * a % kMcg31M;
        e >>= 1;
    }
    return r;
}

// vslNewStream(VSL_BRNG_MCG31, seed) + vslLeapfrogStream(k, n)
Mcg31 mcg31_stream(uint64_t seed, int k, int n)
{
    Mcg31 s;
    uint64_t x0 = seed % kMcg31M;
    if (x0 == 0) x0 = 1;
    s.x = x0 * mcg31_pow(kMcg31A, (uint64_t)k) % kMcg31M;
    s.stride = mcg31_pow(kMcg31A, (uint64_t)n);
    return s;
}

// VSL_RNG_METHOD_GAUSSIAN_BOXMULLER from two consecutive uniforms
double box_muller(Mcg31 &s)
{
    const double u1 = s.next(0.0, 1.0);
    const double u2 = s.next(0.0, 1.0);
    return std::sqrt(-2.0 * std::log(u1)) * std::sin(2.0 * kPi * u2);
}

struct C {
    double re, im;
};

double norm_square(const C *x, int d)
{
    double s = 0.0;
    for (int i = 0; i < d; i++) s += x[i].re * x[i].re + x[i].im * x[i].im;
    return s;
}

}  // namespace

struct qjx_lpn {
    qjx_ctx *c = nullptr;
    qjx_lpn_desc p{};
    int M = 0, half = 0, d = 0, period = 0;
    double max_energy = 0.0;
    std::vector<std::vector<C>> R;                  // random observables, d x d
    std::vector<Mcg31> var;                         // perturbation streams (streams_var)
    std::vector<double> mean, energy, mean_lpn, energy_lpn, lambda, lambda_now, delta_s;
    std::vector<double> norm, disp, m2, mean_start, adr;   // calc_chars_std; adr is M x d
    std::vector<int> num_renorms;
    std::vector<std::vector<C>> obs, obs_lpn;       // [tr][obs]
    std::vector<std::vector<double>> lambdas, deltas_s, deltas_f;
    std::vector<C> psi;                             // host copy of the states, M x d
    std::vector<double> eta;
    std::vector<uint64_t> rng;
    std::vector<C> tmp;

    C *phi(int tr) { return psi.data() + (size_t)tr * d; }

    // mean <k> and energy <H + E H_drv> as calc_chars_std computes them
    // (get_mean_simple, get_energy: experiment.cpp:1395-1469)
    void mean_energy(const C *f, double *mn_out, double *en_out) const
    {
        const double nrm = norm_square(f, d);
        double mn = 0.0;
        for (int k = 0; k < d; k++) mn += double(k) * ((f[k].re * f[k].re + f[k].im * f[k].im) * (1.0 / nrm));
        const double sn = std::sqrt(nrm);
        const double E0 = c->model.prm_E;
        double en = 0.0;
        for (int i = 0; i < d; i++) {
            double sre = 0.0, sim = 0.0;
            for (int j = std::max(0, i - 1); j <= std::min(d - 1, i + 1); j++) {
                const double h = i == j ? c->Hdiag[i] : c->Hoff[std::min(i, j)];
                const double hv = h + E0 * (i == j ? c->Hdrv[i] : 0.0);
                sre += hv * f[j].re / sn - 0.0 * f[j].im / sn;
                sim += hv * f[j].im / sn + 0.0 * f[j].re / sn;
            }
            en += f[i].re / sn * sre + f[i].im / sn * sim;
        }
        *mn_out = mn;
        *en_out = en;
    }

    // get_random_obs: <phi|R|phi> / ||phi||^2 in the reference's loop order
    C random_obs(const C *f, const std::vector<C> &Rm)
    {
        const double n = std::sqrt(norm_square(f, d));
        tmp.resize(2 * (size_t)d);
        C *pn = tmp.data(), *mult = tmp.data() + d;
        for (int i = 0; i < d; i++) pn[i] = {f[i].re / n, f[i].im / n};
        for (int i = 0; i < d; i++) {
            double re = 0.0, im = 0.0;
            for (int j = 0; j < d; j++) {
                const C &a = Rm[(size_t)i * d + j];
                re = re + (a.re * pn[j].re - a.im * pn[j].im);
                im = im + (a.re * pn[j].im + a.im * pn[j].re);
            }
            mult[i] = {re, im};
        }
        double rr = 0.0, ri = 0.0;
        for (int i = 0; i < d; i++) {
            const double cr = pn[i].re, ci = -pn[i].im;
            rr = rr + (cr * mult[i].re - ci * mult[i].im);
            ri = ri + (ci * mult[i].re + cr * mult[i].im);
        }
        return {rr, ri};
    }

    // abs_diag_rho of trajectory tr (both calc_chars_std and calc_chars_lpn write it)
    void write_adr(int tr)
    {
        const C *f = phi(tr);
        const double nrm = norm_square(f, d);
        double *a = adr.data() + (size_t)tr * d;
        for (int k = 0; k < d; k++) a[k] = (f[k].re * f[k].re + f[k].im * f[k].im) * (1.0 / nrm);
    }

    // calc_chars_std(_start) (core.cpp:665-725): norm, mean, dispersion against
    // mean_start, m2, energy, abs_diag_rho, random observables
    void chars_std(int tr, bool is_start = false)
    {
        norm[tr] = norm_square(phi(tr), d);
        write_adr(tr);
        mean_energy(phi(tr), &mean[tr], &energy[tr]);
        if (is_start) mean_start[tr] = mean[tr];
        disp[tr] = (mean[tr] - mean_start[tr]) * (mean[tr] - mean_start[tr]);
        const double *a = adr.data() + (size_t)tr * d;
        double mm2 = 0.0;
        for (int k = 0; k < d; k++) mm2 += (double(k) - mean[tr]) * (double(k) - mean[tr]) * a[k];
        m2[tr] = mm2;
        for (size_t o = 0; o < R.size(); o++) obs[tr][o] = random_obs(phi(tr), R[o]);
    }

    // calc_chars_lpn (core.cpp:761-792)
    void chars_lpn(int tr)
    {
        write_adr(tr);
        mean_energy(phi(tr), &mean_lpn[tr], &energy_lpn[tr]);
        for (size_t o = 0; o < R.size(); o++) obs_lpn[tr][o] = random_obs(phi(tr), R[o]);
    }

    // calc_delta_s (lpn values) / calc_delta_f (std values)
    double delta(int tr, int base, bool from_lpn) const
    {
        const std::vector<double> &mn = from_lpn ? mean_lpn : mean;
        const std::vector<double> &en = from_lpn ? energy_lpn : energy;
        const std::vector<std::vector<C>> &ob = from_lpn ? obs_lpn : obs;
        if (p.lpn_type == -1) return std::fabs(mn[tr] - mn[base]) / double(d);
        if (p.lpn_type == -2) return std::fabs(en[tr] - en[base]) / max_energy;
        const C b = ob[base][p.lpn_type], v = ob[tr][p.lpn_type];
        const double t = std::pow(b.re - v.re, 2) + std::pow(b.im - v.im, 2);
        return std::sqrt(t);
    }

    // var_trajectory_lpn: phi = base + eps var renormalised to the base norm, eps
    // bisected in log10 until the lpn distance is lpn_delta_s within lpn_eps_error
    void perturb(int tr, int base)
    {
        double lo = p.lpn_eps_low, hi = p.lpn_eps_high, diff = hi - lo;
        int deep = 0;
        double eps = std::pow(10.0, (lo + hi) * 0.5);
        C *f = phi(tr);
        const C *orig = phi(base);
        std::vector<double> v(2 * (size_t)d);
        for (int i = 0; i < 2 * d; i++) v[i] = var[tr].next(-1.0, 1.0);
        std::vector<C> w(d);
        for (int i = 0; i < d; i++) w[i] = {v[i], v[d + i]};
        const double nv2 = norm_square(w.data(), d);
        for (int i = 0; i < d; i++) w[i] = {w[i].re / std::sqrt(nv2), w[i].im / std::sqrt(nv2)};
        const double n2 = norm_square(orig, d);
        while (deep < p.lpn_eps_deep) {
            for (int i = 0; i < d; i++) f[i] = {orig[i].re + eps * w[i].re, orig[i].im + eps * w[i].im};
            const double nm2 = norm_square(f, d);
            for (int i = 0; i < d; i++) f[i] = {f[i].re / std::sqrt(nm2), f[i].im / std::sqrt(nm2)};
            for (int i = 0; i < d; i++) f[i] = {f[i].re * std::sqrt(n2), f[i].im * std::sqrt(n2)};
            chars_lpn(tr);
            const double ds = delta(tr, base, true);
            if (std::fabs(ds - p.lpn_delta_s) < p.lpn_eps_error) break;
            if (ds > p.lpn_delta_s) {
                hi -= 0.5 * diff;
                diff = hi - lo;
                deep++;
                eps = std::pow(10.0, (lo + hi) * 0.5);
            }
            if (ds < p.lpn_delta_s) {
                lo += 0.5 * diff;
                diff = hi - lo;
                deep++;
                eps = std::pow(10.0, (lo + hi) * 0.5);
            }
            if (!(ds > p.lpn_delta_s) && !(ds < p.lpn_delta_s)) break;   // equal or NaN
        }
    }

    void copy(int tr, int base, bool state)
    {
        rng[tr] = rng[base];
        eta[tr] = eta[base];
        if (state) std::memcpy(phi(tr), phi(base), sizeof(C) * d);
    }

    // calc_chars_lpn_start
    void start(int tr, int base)
    {
        lambda[tr] = 0.0;
        lambda_now[tr] = 0.0;
        mean_lpn[tr] = mean[tr];
        energy_lpn[tr] = energy[tr];
        delta_s[tr] = delta(tr, base, false);
        obs_lpn[tr] = obs[tr];
    }

    int pull()
    {
        return qjx_get_state(c, reinterpret_cast<double *>(psi.data()), eta.data(), nullptr, nullptr, rng.data());
    }

    int push()
    {
        return qjx_set_state(c, reinterpret_cast<const double *>(psi.data()), eta.data(), nullptr, rng.data());
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
    const int d = c->d, M = c->M;
    for (int i = 0; i < d; i++) L->max_energy = std::max(L->max_energy, std::fabs(c->Hdiag[i]));   // init_obs_lpn
    // init_random_obs: X = (G_re + i G_im) / 2 from d^2 + d^2 Gaussians, R = X X^dag,
    // normalised by Tr R (type 1) or Tr R / d (type 2)
    const size_t dd = (size_t)d * d;
    L->R.resize(desc->num_random_obs);
    for (int o = 0; o < desc->num_random_obs; o++) {
        Mcg31 s = mcg31_stream(3873247ULL, desc->random_obs_seed + o, desc->random_obs_mns);
        std::vector<double> gr(dd), gi(dd), xr(dd), xi(dd);
        for (size_t i = 0; i < dd; i++) gr[i] = box_muller(s);
        for (size_t i = 0; i < dd; i++) gi[i] = box_muller(s);
        for (size_t i = 0; i < dd; i++) {
            xr[i] = 0.5 * gr[i];
            xi[i] = 0.5 * gi[i];
        }
        std::vector<C> &Rm = L->R[o];
        Rm.assign(dd, C{0.0, 0.0});
        for (int i = 0; i < d; i++)
            for (int j = 0; j < d; j++) {
                double re = 0.0, im = 0.0;
                for (int k = 0; k < d; k++) {
                    const double ar = xr[(size_t)i * d + k], ai = xi[(size_t)i * d + k];
                    const double br = xr[(size_t)j * d + k], bi = -xi[(size_t)j * d + k];
                    re = re + (ar * br - ai * bi);
                    im = im + (ar * bi + ai * br);
                }
                Rm[(size_t)i * d + j] = {re, im};
            }
        if (desc->random_obs_type > 0) {
            double tr = 0.0;
            for (int i = 0; i < d; i++) tr += Rm[(size_t)i * d + i].re;
            if (desc->random_obs_type == 2) tr = tr / double(d);
            for (size_t i = 0; i < dd; i++) Rm[i] = {Rm[i].re / tr, Rm[i].im / tr};
        }
    }
    L->var.resize(M);
    for (int tr = 0; tr < M; tr++) L->var[tr] = mcg31_stream(777, tr, M);
    L->mean.assign(M, 0.0);
    L->energy.assign(M, 0.0);
    L->mean_lpn.assign(M, 0.0);
    L->energy_lpn.assign(M, 0.0);
    L->lambda.assign(M, 0.0);
    L->lambda_now.assign(M, 0.0);
    L->delta_s.assign(M, 0.0);
    L->num_renorms.assign(M, 0);
    L->norm.assign(M, 0.0);
    L->disp.assign(M, 0.0);
    L->m2.assign(M, 0.0);
    L->mean_start.assign(M, 0.0);
    L->adr.assign((size_t)M * d, 0.0);
    L->obs.assign(M, std::vector<C>(desc->num_random_obs, C{0.0, 0.0}));
    L->obs_lpn.assign(M, std::vector<C>(desc->num_random_obs, C{0.0, 0.0}));
    L->lambdas.assign(M, {});
    L->deltas_s.assign(M, {});
    L->deltas_f.assign(M, {});
    L->psi.assign((size_t)M * d, C{0.0, 0.0});
    L->eta.assign(M, 0.0);
    L->rng.assign(M, 0);
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
    if ((rc = L->pull())) return rc;
    for (int tr = 0; tr < L->half; tr++) {
        L->chars_std(tr, true);
        L->start(tr, tr);
    }
    for (int tr = 0; tr < L->M; tr++) {
        const int base = tr >= L->half ? tr - L->half : tr;
        if (tr >= L->half) {
            L->copy(tr, base, true);
            L->perturb(tr, base);
        }
        L->chars_std(tr, true);
        L->start(tr, base);
    }
    if ((rc = L->push())) return rc;
    L->period = 0;
    return QJX_OK;
}

// n observation periods (obser_process's period loop): all trajectories advance one
// period on the device, then the pairs are measured and, past a threshold (or every
// period with lambda_per_periods), renormalised.
extern "C" int qjx_lpn_run_periods(qjx_lpn *L, int n)
{
    if (!L) return lpn_fail(QJX_ERR_ARG, "null LPN handle");
    if (n < 0) return lpn_fail(QJX_ERR_ARG, "negative period count");
    for (int k = 0; k < n; k++) {
        int rc = qjx_run_periods(L->c, 1, 1);
        if (rc) return rc;
        if ((rc = L->pull())) return rc;
        for (int tr = 0; tr < L->M; tr++) L->chars_std(tr);
        for (int tr = 0; tr < L->half; tr++) L->chars_lpn(tr);
        const double curr_time = double(L->period + 1) * L->c->T;
        bool changed = false;
        for (int tr = L->half; tr < L->M; tr++) {
            const int base = tr - L->half;
            const double df = L->delta(tr, base, false);
            const bool renorm = L->p.lambda_per_periods != 0 ||
                                df > std::max(L->p.lpn_delta_f_high, L->delta_s[tr] * 10.0) ||
                                df < L->p.lpn_delta_f_low;
            if (renorm) {
                L->lambda[tr] += std::log(df / L->delta_s[tr] + 1.0e-16);
                L->num_renorms[tr] += 1;
                if (L->p.save_lambdas > 0) {
                    L->lambdas[tr].push_back(df / L->delta_s[tr]);
                    L->deltas_s[tr].push_back(L->delta_s[tr]);
                    L->deltas_f[tr].push_back(df);
                }
                L->copy(tr, base, false);
                L->perturb(tr, base);
                L->chars_lpn(tr);
                L->delta_s[tr] = L->delta(tr, base, true);
                L->lambda_now[tr] = L->lambda[tr] / curr_time;
                changed = true;
            } else {
                L->chars_lpn(tr);
                L->lambda_now[tr] = (L->lambda[tr] + std::log(df / L->delta_s[tr] + 1.0e-16)) / curr_time;
            }
        }
        if (changed && (rc = L->push())) return rc;
        L->period++;
    }
    return QJX_OK;
}

extern "C" int qjx_lpn_get(qjx_lpn *L, double *lambda_now, double *lambda, int *num_renorms, double *mean_lpn,
                           double *energy_lpn, double *delta_s)
{
    if (!L) return lpn_fail(QJX_ERR_ARG, "null LPN handle");
    for (int tr = 0; tr < L->M; tr++) {
        if (lambda_now) lambda_now[tr] = L->lambda_now[tr];
        if (lambda) lambda[tr] = L->lambda[tr];
        if (num_renorms) num_renorms[tr] = L->num_renorms[tr];
        if (mean_lpn) mean_lpn[tr] = L->mean_lpn[tr];
        if (energy_lpn) energy_lpn[tr] = L->energy_lpn[tr];
        if (delta_s) delta_s[tr] = L->delta_s[tr];
    }
    return QJX_OK;
}

extern "C" int qjx_lpn_random_obs(qjx_lpn *L, double *obs, double *obs_lpn)
{
    if (!L) return lpn_fail(QJX_ERR_ARG, "null LPN handle");
    const size_t no = L->R.size();
    for (int tr = 0; tr < L->M; tr++)
        for (size_t o = 0; o < no; o++) {
            const size_t i = 2 * ((size_t)tr * no + o);
            if (obs) {
                obs[i] = L->obs[tr][o].re;
                obs[i + 1] = L->obs[tr][o].im;
            }
            if (obs_lpn) {
                obs_lpn[i] = L->obs_lpn[tr][o].re;
                obs_lpn[i + 1] = L->obs_lpn[tr][o].im;
            }
        }
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
    if (obs < 0 || obs >= (int)L->R.size()) return lpn_fail(QJX_ERR_ARG, "random observable out of range");
    std::memcpy(R, L->R[obs].data(), sizeof(C) * L->R[obs].size());
    return QJX_OK;
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
    for (int tr = 0; tr < M; tr++) {
        if (norm) norm[tr] = L->norm[tr];
        if (mean) mean[tr] = L->mean[tr];
        if (dispersion) dispersion[tr] = L->disp[tr];
        if (m2) m2[tr] = L->m2[tr];
        if (energy) energy[tr] = L->energy[tr];
    }
    if (adr) std::memcpy(adr, L->adr.data(), sizeof(double) * L->adr.size());
    if (adr_avg) {
        const double n = double(end - begin);
        for (int k = 0; k < d; k++) adr_avg[k] = 0.0;
        for (int tr = begin; tr < end; tr++)
            for (int k = 0; k < d; k++) adr_avg[k] += L->adr[(size_t)tr * d + k] / n;
    }
    return QJX_OK;
}
