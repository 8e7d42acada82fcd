// qjx_config.cpp -- the drop-in runner: the reference's config.txt / params.txt in,
// its STD-experiment output files out, with the propagation on the GPU engine.
//
// Restates, for sys_id 0 / task_id 1, 4 and 7 (STD, STD deep, LPN) / prop_id 0 and 1 (reference files under
// source/cpp/QJX/QJX/):
//   init_params / set_param (config.cpp:29-145), the single-run main
//   (main.cpp:3-192) and Processor::process (processor.cpp:50-75),
//   DimerOutputBehavior::suffix_param (outputter.cpp:3-44, 353-388),
//   init_dump_periods (newdel_exp.cpp:408-455),
//   StdExperimentBehaviour::trans_process / obser_process (experiment.cpp:411-542),
//   StdDeepExperimentBehaviour (experiment.cpp:769-862), init_dump_periods_deep
//   (newdel_exp.cpp:377-406),
//   DimerCoreBehaviour::dump_std / dump_std_evo (core.cpp:918-1084),
//   dump_adr_single / dump_adr_avg (dump.cpp:3-61), dump_phi(_evo) (core.cpp:12567-12613),
//   the text writers (utils.cpp:64-467).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/qjx.h"

namespace {

thread_local std::string g_cfg_err;

struct RunCfg {
    int sys_id = 0, task_id = 0, prop_id = 0, is_debug = 0, is_pp = 0;
    std::string init_fn, path;
    int num_threads = 1;
    int num_trajectories = 1, seed = 0, mns = 1000000;
    int num_tp_periods = 0, num_obs_periods = 10, ex_deep = 16, rk_ns = 1000;
    std::map<std::string, double> params;
};

std::vector<std::string> split_tokens(const std::string &str, const std::string &delim)
{
    // config.cpp:3-27: empty tokens are dropped
    std::vector<std::string> tokens;
    size_t prev = 0, pos = 0;
    do {
        pos = str.find(delim, prev);
        if (pos == std::string::npos) pos = str.length();
        std::string token = str.substr(prev, pos - prev);
        if (!token.empty()) tokens.push_back(token);
        prev = pos + delim.length();
    } while (pos < str.length() && prev < str.length());
    return tokens;
}

void set_param(RunCfg &c, const std::string &k, const std::string &v)
{
    const int iv = std::atoi(v.c_str());
    if (k == "sys_id") c.sys_id = iv;
    else if (k == "task_id") c.task_id = iv;
    else if (k == "prop_id") c.prop_id = iv;
    else if (k == "is_debug") c.is_debug = iv;
    else if (k == "is_pp") c.is_pp = iv;
    else if (k == "init_fn") c.init_fn = v;
    else if (k == "path") c.path = v;
    else if (k == "num_threads") c.num_threads = iv;
    else if (k == "num_trajectories") c.num_trajectories = iv;
    else if (k == "seed") c.seed = iv;
    else if (k == "mns") c.mns = iv;
    else if (k == "num_tp_periods") c.num_tp_periods = iv;
    else if (k == "num_obs_periods") c.num_obs_periods = iv;
    else if (k == "ex_deep") c.ex_deep = iv;
    else if (k == "rk_ns") c.rk_ns = iv;
}

bool read_files(RunCfg &c, const char *fn_config, const char *fn_param)
{
    std::string line;
    std::ifstream cf(fn_config);
    if (!cf.is_open()) {
        g_cfg_err = std::string("unable to open config file ") + fn_config;
        return false;
    }
    while (std::getline(cf, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        std::vector<std::string> t = split_tokens(line, " ");
        if (t.size() == 2) set_param(c, t[0], t[1]);
    }
    std::ifstream pf(fn_param);
    if (!pf.is_open()) {
        g_cfg_err = std::string("unable to open params file ") + fn_param;
        return false;
    }
    std::string key;
    double val = 0.0;
    while (std::getline(pf, line)) {
        std::istringstream iss(line);
        if (iss >> key >> val) c.params[key] = val;
    }
    return true;
}

// The reference looks every key up with params.find(key)->second, which is
// undefined for a missing key; here a missing key is an error naming it.
struct Params {
    const std::map<std::string, double> &m;
    std::string missing;
    double get(const std::string &k)
    {
        auto it = m.find(k);
        if (it == m.end()) {
            if (missing.empty()) missing = k;
            return 0.0;
        }
        return it->second;
    }
};

std::string suffix_param(const RunCfg &c, Params &p)
{
    const int prec = 4;   // Processor::process passes 4 (processor.cpp:52)
    std::stringstream s;
    s << "_setup(" << c.sys_id << "_" << c.task_id << "_" << c.prop_id << ")";
    s << "_rnd(" << c.seed << "_" << c.mns << ")";
    s << "_N(" << int(p.get("N")) << ")";
    s << "_diss(" << int(p.get("diss_type")) << "_" << std::setprecision(prec) << std::fixed
      << p.get("diss_gamma") << "_" << std::setprecision(prec) << std::fixed << p.get("diss_phase") << ")";
    s << "_drv(" << int(p.get("dimer_drv_type")) << "_" << std::setprecision(prec) << std::fixed
      << p.get("dimer_drv_ampl") << "_" << std::setprecision(prec) << std::fixed << p.get("dimer_drv_freq")
      << "_" << std::setprecision(prec) << std::fixed << p.get("dimer_drv_phase") << ")";
    s << "_prm(" << std::setprecision(prec) << std::fixed << p.get("dimer_prm_E") << "_"
      << std::setprecision(prec) << std::fixed << p.get("dimer_prm_U") << "_" << std::setprecision(prec)
      << std::fixed << p.get("dimer_prm_J") << ")";
    s << "_start(" << int(p.get("start_type")) << "_" << int(p.get("start_state")) << ")";
    s << ".txt";
    return s.str();
}

std::vector<int> dump_periods(int dump_type, int dump_num, int num_obs_periods)
{
    std::vector<int> dp;
    if (dump_type == 0) {
        dp.assign(dump_num + 1, 0);
        const int shift = num_obs_periods / dump_num;
        for (int i = 0; i < dump_num; i++) dp[i + 1] = (i + 1) * shift;
    } else if (dump_type == 1) {
        dp.assign(dump_num + 2, 0);   // entry 0 is left unset by the reference; 0 here
        const double begin_decade = std::log10(1.0);
        const double end_decade = std::log10(double(num_obs_periods));
        const double num_decades = end_decade - begin_decade;
        const double num_decades_dump = double(dump_num) / num_decades;
        for (int i = 0; i < dump_num + 1; i++) {
            const int cur = int(std::pow(10.0, begin_decade) * std::pow(10.0, (1.0 / num_decades_dump) * double(i)));
            dp[i + 1] = cur > dp[i] ? cur : dp[i] + 1;
        }
    }
    return dp;
}

// utils.cpp writers: setprecision(16) scientific, one value per line.
bool save_doubles(const std::string &fn, const double *v, size_t n, bool append)
{
    std::ofstream o(fn, append ? std::ios::app : std::ios::out);
    if (!o.is_open()) return false;
    o << std::setprecision(16) << std::scientific;
    for (size_t i = 0; i < n; i++) o << v[i] << std::endl;
    return true;
}

bool save_ints(const std::string &fn, const int *v, size_t n, bool append)
{
    std::ofstream o(fn, append ? std::ios::app : std::ios::out);
    if (!o.is_open()) return false;
    for (size_t i = 0; i < n; i++) o << v[i] << std::endl;
    return true;
}

// save_2d_inv_double_data (utils.cpp:199-226): rows = dumps, columns = trajectories.
bool save_2d_inv(const std::string &fn, const std::vector<double> &v, int rows, int cols)
{
    std::ofstream o(fn);
    if (!o.is_open()) return false;
    o << std::setprecision(16) << std::scientific;
    for (int r = 0; r < rows; r++) {
        for (int c = 0; c < cols; c++) o << v[(size_t)c * rows + r] << " ";
        o << std::endl;
    }
    return true;
}

bool save_complex(const std::string &fn, const double *v, size_t n, bool append)
{
    std::ofstream o(fn, append ? std::ios::app : std::ios::out);
    if (!o.is_open()) return false;
    o << std::setprecision(16) << std::scientific;
    for (size_t i = 0; i < n; i++) o << v[2 * i] << " " << v[2 * i + 1] << std::endl;
    return true;
}

// save_2d_vector<std::complex<double>> (utils.h:53-71): rows = dumps, columns =
// trajectories, each entry as std::complex prints it, "(re,im)".
bool save_2d_complex(const std::string &fn, const std::vector<std::vector<double>> &v)
{
    std::ofstream o(fn);
    if (!o.is_open()) return false;
    o << std::setprecision(16) << std::scientific;
    const size_t cols = v.size(), rows = cols ? v[0].size() / 2 : 0;
    for (size_t r = 0; r < rows; r++)
        for (size_t c = 0; c < cols; c++)
            o << "(" << v[c][2 * r] << "," << v[c][2 * r + 1] << ")" << (c + 1 < cols ? "\t" : "\n");
    return true;
}

// LpnMultNewDelBehaviour / LpnMultExperimentBehaviour (newdel_exp.cpp:29-48,
// experiment.cpp:137-286) with dump_std / dump_lpn / dump_std_evo / dump_lpn_evo
// (core.cpp:918-1130) and dump_adr_single / dump_adr_avg(_mult) (dump.cpp:3-98) on an
// initialised context.  Returns QJX_OK or an error (message in g_cfg_err).
int run_lpn(const RunCfg &c, Params &p, qjx_ctx *ctx, const std::string &sfx, const std::vector<int> &dps, int M,
            int d)
{
    const std::string &path = c.path;
    const int dnt = (int)dps.size();
    qjx_lpn_desc ld{};
    ld.lpn_type = int(p.get("lpn_type"));
    ld.lpn_eps_deep = int(p.get("lpn_eps_deep"));
    ld.lpn_eps_error = p.get("lpn_eps_error");
    ld.lpn_eps_high = p.get("lpn_eps_high");
    ld.lpn_eps_low = p.get("lpn_eps_low");
    ld.lpn_delta_s = p.get("lpn_delta_s");
    ld.lpn_delta_f_high = p.get("lpn_delta_f_high");
    ld.lpn_delta_f_low = p.get("lpn_delta_f_low");
    ld.lambda_per_periods = int(p.get("lambda_per_periods"));
    ld.save_lambdas = int(p.get("save_lambdas"));
    ld.num_random_obs = int(p.get("num_random_obs"));
    ld.random_obs_seed = int(p.get("random_obs_seed"));
    ld.random_obs_mns = int(p.get("random_obs_mns"));
    ld.random_obs_type = int(p.get("random_obs_type"));
    const int jump = int(p.get("jump"));
    const int dump_obs = int(p.get("dump_obs"));
    const int dump_evo_sep = int(p.get("dump_evo_sep"));
    const int dump_evo_avg = int(p.get("dump_evo_avg"));
    const int dump_adr_sep = int(p.get("dump_adr_sep"));
    const int dump_adr_avg = int(p.get("dump_adr_avg"));
    if (!p.missing.empty()) {
        g_cfg_err = "params.txt lacks key " + p.missing;
        return QJX_ERR_ARG;
    }
    qjx_lpn *L = nullptr;
    int rc = qjx_lpn_create(ctx, &ld, &L);
    if (rc < 0) {
        g_cfg_err = qjx_last_error();
        return rc;
    }
    const int no = ld.num_random_obs, half = M / 2;
    std::vector<double> norm(M), mean(M), disp(M), m2(M), energy(M), mean_start(M), adr((size_t)M * d), avg(d);
    std::vector<double> lambda_now(M), mean_lpn(M), energy_lpn(M);
    std::vector<int> renorms(M);
    std::vector<double> norm_evo((size_t)M * dnt, 0.0), mean_evo = norm_evo, disp_evo = norm_evo, m2_evo = norm_evo,
                        energy_evo = norm_evo, lambda_evo = norm_evo, mean_lpn_evo = norm_evo,
                        energy_lpn_evo = norm_evo;
    // random observables per obs and trajectory: (re, im) per dump
    std::vector<std::vector<std::vector<double>>> ro_evo(no, std::vector<std::vector<double>>(M)),
        ro_lpn_evo = ro_evo;
    std::vector<double> ro((size_t)2 * M * std::max(no, 1)), ro_lpn = ro;
    auto fail = [&](int code, const std::string &msg) {
        g_cfg_err = msg;
        qjx_lpn_destroy(L);
        return code;
    };
    // evo_chars_std / evo_chars_lpn at dump_id, then dump_adr_single
    auto record = [&](int dump_id, bool append_adr) -> int {
        int r = qjx_lpn_chars(L, norm.data(), mean.data(), disp.data(), m2.data(), energy.data(), adr.data(), 0, 0,
                              nullptr);
        if (r == QJX_OK) r = qjx_lpn_get(L, lambda_now.data(), nullptr, renorms.data(), mean_lpn.data(),
                                        energy_lpn.data(), nullptr);
        if (r == QJX_OK) r = qjx_lpn_random_obs(L, ro.data(), ro_lpn.data());
        if (r < 0) return r;
        for (int tr = 0; tr < M; tr++) {
            const size_t i = (size_t)tr * dnt + dump_id;
            norm_evo[i] = norm[tr];
            mean_evo[i] = mean[tr];
            disp_evo[i] = disp[tr];
            m2_evo[i] = m2[tr];
            energy_evo[i] = energy[tr];
            lambda_evo[i] = lambda_now[tr];
            mean_lpn_evo[i] = mean_lpn[tr];
            energy_lpn_evo[i] = energy_lpn[tr];
            for (int o = 0; o < no; o++) {
                const size_t k = 2 * ((size_t)tr * no + o);
                ro_evo[o][tr].push_back(ro[k]);
                ro_evo[o][tr].push_back(ro[k + 1]);
                ro_lpn_evo[o][tr].push_back(ro_lpn[k]);
                ro_lpn_evo[o][tr].push_back(ro_lpn[k + 1]);
            }
            if (dump_evo_sep == 1 && dump_adr_sep == 1 &&
                !save_doubles(path + "adr_" + std::to_string(tr) + sfx, &adr[(size_t)tr * d], d, append_adr))
                return QJX_ERR_IO;
        }
        return QJX_OK;
    };
    auto adr_avg_dump = [&](int begin, int end, bool append) -> int {
        if (dump_adr_avg != 1) return QJX_OK;
        int r = qjx_lpn_chars(L, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, begin, end, avg.data());
        if (r < 0) return r;
        return save_doubles(path + "adr_avg" + sfx, avg.data(), d, append) ? QJX_OK : QJX_ERR_IO;
    };

    // trans_process (experiment.cpp:137-190)
    if ((rc = qjx_lpn_trans(L, c.num_tp_periods)) < 0) return fail(rc, qjx_last_error());
    if ((rc = record(0, false)) != QJX_OK) return fail(rc, rc == QJX_ERR_IO ? "unable to write output file" : qjx_last_error());
    mean_start = mean;
    if (dump_evo_avg == 1 && (rc = adr_avg_dump(0, M, false)) != QJX_OK) return fail(rc, "unable to write output file");
    // obser_process (experiment.cpp:193-286): every period advances all trajectories
    // and updates the pairs; observables are recorded at the dump periods
    for (int dump_id = 1; dump_id < dnt; dump_id++) {
        if (c.is_pp == 1) std::cout << "dump_id: " << dump_id << std::endl;
        if ((rc = qjx_lpn_run_periods(L, dps[dump_id] - dps[dump_id - 1])) < 0) return fail(rc, qjx_last_error());
        if ((rc = record(dump_id, true)) != QJX_OK)
            return fail(rc, rc == QJX_ERR_IO ? "unable to write output file" : qjx_last_error());
        if (dump_evo_avg == 1 && (rc = adr_avg_dump(0, half, true)) != QJX_OK)
            return fail(rc, "unable to write output file");
    }
    bool ok = true;
    if (dump_obs == 1) {                     // dump_std (core.cpp:918-957) and dump_lpn (core.cpp:959-1004)
        ok = ok && save_doubles(path + "norm" + sfx, norm.data(), M, false);
        ok = ok && save_doubles(path + "mean_start" + sfx, mean_start.data(), M, false);
        ok = ok && save_doubles(path + "mean" + sfx, mean.data(), M, false);
        ok = ok && save_doubles(path + "dispersion" + sfx, disp.data(), M, false);
        ok = ok && save_doubles(path + "m2" + sfx, m2.data(), M, false);
        ok = ok && save_doubles(path + "energy" + sfx, energy.data(), M, false);
        ok = ok && save_doubles(path + "lambda" + sfx, lambda_now.data(), M, false);
        ok = ok && save_ints(path + "num_renorms" + sfx, renorms.data(), M, false);
        ok = ok && save_doubles(path + "mean_lpn" + sfx, mean_lpn.data(), M, false);
        ok = ok && save_doubles(path + "energy_lpn" + sfx, energy_lpn.data(), M, false);
        if (ld.save_lambdas > 0)
            for (int tr = 0; tr < M && ok; tr++) {
                const int n = qjx_lpn_lambdas(L, tr, nullptr, nullptr, nullptr, 0);
                std::vector<double> r(n), ds(n), df(n);
                qjx_lpn_lambdas(L, tr, r.data(), ds.data(), df.data(), n);
                ok = ok && save_doubles(path + "delta_f_by_delta_s_" + std::to_string(tr) + sfx, r.data(), n, false);
                ok = ok && save_doubles(path + "deltas_s_" + std::to_string(tr) + sfx, ds.data(), n, false);
                ok = ok && save_doubles(path + "deltas_f_" + std::to_string(tr) + sfx, df.data(), n, false);
            }
    }
    if (ok && dump_evo_sep == 1 && dump_obs == 1) {   // dump_std_evo and dump_lpn_evo (core.cpp:1008-1130)
        ok = ok && save_ints(path + "periods" + sfx, dps.data(), dnt, false);
        ok = ok && save_2d_inv(path + "norm_evo" + sfx, norm_evo, dnt, M);
        ok = ok && save_2d_inv(path + "mean_evo" + sfx, mean_evo, dnt, M);
        ok = ok && save_2d_inv(path + "dispersion_evo" + sfx, disp_evo, dnt, M);
        ok = ok && save_2d_inv(path + "m2_evo" + sfx, m2_evo, dnt, M);
        ok = ok && save_2d_inv(path + "energy_evo" + sfx, energy_evo, dnt, M);
        for (int o = 0; o < no && ok; o++)
            ok = save_2d_complex(path + "random_obs_evo_" + std::to_string(ld.random_obs_seed + o) + sfx, ro_evo[o]);
        if (jump > 0)
            for (int tr = 0; tr < M && ok; tr++) {
                const int n = qjx_get_jumps(ctx, tr, nullptr, nullptr, nullptr, nullptr, 0);
                std::vector<double> t(n), nr(n), et(n);
                std::vector<int> ty(n);
                qjx_get_jumps(ctx, tr, t.data(), ty.data(), nr.data(), et.data(), n);
                ok = ok && save_doubles(path + "jump_times_" + std::to_string(tr) + sfx, t.data(), n, false);
                ok = ok && save_ints(path + "diss_types_" + std::to_string(tr) + sfx, ty.data(), n, false);
                if (jump > 1) {
                    ok = ok && save_doubles(path + "jump_norms_" + std::to_string(tr) + sfx, nr.data(), n, false);
                    ok = ok && save_doubles(path + "jump_etas_" + std::to_string(tr) + sfx, et.data(), n, false);
                }
            }
        ok = ok && save_2d_inv(path + "lambda_evo" + sfx, lambda_evo, dnt, M);
        ok = ok && save_2d_inv(path + "mean_lpn_evo" + sfx, mean_lpn_evo, dnt, M);
        ok = ok && save_2d_inv(path + "energy_lpn_evo" + sfx, energy_lpn_evo, dnt, M);
        for (int o = 0; o < no && ok; o++)
            ok = save_2d_complex(path + "random_obs_lpn_evo_" + std::to_string(ld.random_obs_seed + o) + sfx,
                                 ro_lpn_evo[o]);
    }
    if (ok && dump_evo_avg == 0) ok = adr_avg_dump(0, half, true) == QJX_OK;
    if (!ok) return fail(QJX_ERR_IO, "unable to write output file");
    qjx_lpn_destroy(L);
    return QJX_OK;
}

}  // namespace

#define CFG_FAIL(code, msg)          \
    do {                             \
        g_cfg_err = (msg);           \
        rc = (code);                 \
        goto done;                   \
    } while (0)

#define IO_CHECK(expr)                                                  \
    do {                                                                \
        if (!(expr)) CFG_FAIL(QJX_ERR_IO, "unable to write output file"); \
    } while (0)

#define API_CHECK(expr)                                  \
    do {                                                 \
        int r_ = (expr);                                 \
        if (r_ < 0) CFG_FAIL(r_, qjx_last_error());      \
    } while (0)

namespace qjx {
void set_last_error(const std::string &msg);
}

extern "C" int qjx_run_config(const char *config_path, const char *params_path, int precision,
                              int steps_per_period)
{
    int rc = QJX_OK;
    RunCfg c;
    qjx_ctx *ctx = nullptr;
    std::string sfx;
    int M = 0, d = 0, jump = 0, dump_obs = 0, dump_evo_sep = 0, dump_evo_avg = 0, dump_adr_sep = 0,
        dump_adr_avg = 0, dump_phi = 0, dump_phi_evo = 0, jumps_lim = 0, dnt = 0;
    bool deep = false;
    int deep_steps = 0, sub_steps = 0;   // StdDeep: sub-periods per half period, steps per sub-period
    int ro_num = 0, ro_seed = 0, ro_mns = 0, ro_type = 0;   // random observables of the STD path
    std::vector<std::vector<std::vector<double>>> ro_evo;   // [obs][trajectory][(re, im) per dump]
    std::vector<double> ro;
    std::vector<int> dps;
    std::vector<double> norm_evo, mean_evo, disp_evo, m2_evo, energy_evo, adr, adr_avg, psi;
    std::vector<double> norm, mean_start, mean, disp, m2, energy;
    std::vector<int> nj;
    auto t0 = std::chrono::steady_clock::now();

    if (!config_path || !params_path) {
        qjx::set_last_error("null path");
        return QJX_ERR_ARG;
    }
    if (!read_files(c, config_path, params_path)) {
        qjx::set_last_error(g_cfg_err);
        return QJX_ERR_IO;
    }
    {
        Params p{c.params, ""};
        if (c.sys_id != 0 || (c.task_id != 1 && c.task_id != 4 && c.task_id != 7) ||
            (c.prop_id != 0 && c.prop_id != 1))
            CFG_FAIL(QJX_ERR_ARG, "only sys_id 0 (dimer), task_id 1 (STD), 4 (STD deep) or 7 (LPN), prop_id 0 "
                                  "(QJ) or 1 (RK) are implemented");
        deep = c.task_id == 4;
        sfx = suffix_param(c, p);
        qjx_model_desc md{};
        md.N = int(p.get("N"));
        md.prm_E = p.get("dimer_prm_E");
        md.prm_U = p.get("dimer_prm_U");
        md.prm_J = p.get("dimer_prm_J");
        md.drv_ampl = p.get("dimer_drv_ampl");
        md.drv_freq = p.get("dimer_drv_freq");
        md.diss_gamma = p.get("diss_gamma");
        md.ex_deep = c.ex_deep;
        md.steps_per_period = steps_per_period > 0 ? steps_per_period : 4;
        md.precision = precision;
        jump = int(p.get("jump"));
        jumps_lim = int(p.get("jumps_counts"));
        dump_obs = int(p.get("dump_obs"));
        dump_evo_sep = int(p.get("dump_evo_sep"));
        dump_evo_avg = int(p.get("dump_evo_avg"));
        dump_adr_sep = int(p.get("dump_adr_sep"));
        dump_adr_avg = int(p.get("dump_adr_avg"));
        dump_phi = int(p.get("dump_phi"));
        dump_phi_evo = int(p.get("dump_phi_evo"));
        const int dump_type = int(p.get("dump_type"));
        const int dump_num = int(p.get("dump_num"));
        const int num_random_obs = int(p.get("num_random_obs"));
        const int start_type = int(p.get("start_type"));
        const int start_state = int(p.get("start_state"));
        const int drv_type = int(p.get("dimer_drv_type"));
        const double drv_phase = p.get("dimer_drv_phase");
        if (deep) deep_steps = int(p.get("deep_num_steps"));
        if (!p.missing.empty()) CFG_FAIL(QJX_ERR_ARG, "params.txt lacks key " + p.missing);
        // random observables (init_random_obs, newdel.cpp:3652-3722) are read only
        // when there are any, as the reference reads their keys
        if (num_random_obs < 0) CFG_FAIL(QJX_ERR_ARG, "num_random_obs must be >= 0");
        if (num_random_obs > 0 && c.task_id != 7) {
            ro_seed = int(p.get("random_obs_seed"));
            ro_mns = int(p.get("random_obs_mns"));
            ro_type = int(p.get("random_obs_type"));
            if (!p.missing.empty()) CFG_FAIL(QJX_ERR_ARG, "params.txt lacks key " + p.missing);
            ro_num = num_random_obs;
        }
        if (deep) {
            // StdDeep (task 4): a dump after each of the 2 deep_num_steps sub-periods of
            // every period; dump_type / dump_num do not apply (init_dump_periods_deep,
            // newdel_exp.cpp:377-406).  The QJ tree's branch head is T / (2 deep_num_steps)
            // with two children (init_split_structure_dimer_deep, core.cpp:9928-9981); each
            // child halves that dt twice before its first level (init_split_branches,
            // qj_proc.cpp:15-24), so level 0 is 2 steps of T / (8 deep_num_steps) per child
            // and a sub-period runs both children (one_sub_period_deep, qj_proc.cpp:345-360):
            // 4 level-0 steps per sub-period, 8 deep_num_steps per period.  The RK step is
            // T / (rk_ns 2 deep_num_steps) (init_rk_deep, rk_proc.cpp:65-76).
            if (deep_steps < 1) CFG_FAIL(QJX_ERR_ARG, "deep_num_steps must be >= 1");
            const int sub = 2 * deep_steps;
            dps.assign((size_t)sub * c.num_obs_periods + 1, 0);
            for (int pid = 0; pid < c.num_obs_periods; pid++)
                for (int k = 0; k < sub; k++) dps[(size_t)pid * sub + k + 1] = pid;
            md.steps_per_period = 8 * deep_steps;
            sub_steps = c.prop_id == 1 ? c.rk_ns : 4;
        } else {
            if (dump_type != 0 && dump_type != 1) CFG_FAIL(QJX_ERR_ARG, "dump_type must be 0 or 1");
            if (dump_num < 1 || (dump_type == 0 && c.num_obs_periods < dump_num))
                CFG_FAIL(QJX_ERR_ARG, "dump_num must be in [1, num_obs_periods]");
            dps = dump_periods(dump_type, dump_num, c.num_obs_periods);
        }
        dnt = (int)dps.size();

        if (c.prop_id == 1) {
            // RKPropagateBehavior (propagator.cpp:66-129): complex128 as the reference,
            // whatever GEMM form was asked for the quantum-jump path
            qjx_rk_desc rk{};
            rk.rk_ns = deep ? c.rk_ns * 2 * deep_steps : c.rk_ns;
            rk.drv_type = drv_type;
            rk.drv_phase = drv_phase;
            md.precision = QJX_FP64;
            API_CHECK(qjx_ctx_create_rk(&md, &rk, &ctx));
        } else {
            API_CHECK(qjx_ctx_create(&md, &ctx));
        }
        M = c.num_trajectories;
        d = md.N + 1;
        qjx_ensemble_desc ed{};
        ed.num_trajectories = M;
        ed.seed = c.seed;
        ed.mns = c.mns;
        ed.start_type = start_type;
        ed.start_state = start_state;
        ed.jump = jump;
        // Logged jumps are drained after every period: the jump rate is at most
        // gamma (N + 2), so 1.5 gamma (N + 2) T + 64 records per period suffice.
        const double T = 2.0 * 3.1415926535897932384626433832795 / md.drv_freq;
        ed.log_capacity = jump > 0 ? int(std::ceil(1.5 * md.diss_gamma * (md.N + 2) * T)) + 64 : 0;
        API_CHECK(qjx_ensemble_init(ctx, &ed));
        if (c.task_id == 7) {
            rc = run_lpn(c, p, ctx, sfx, dps, M, d);
            goto done;
        }

        const std::string &path = c.path;
        norm.assign(M, 0.0);
        mean_start.assign(M, 0.0);
        mean.assign(M, 0.0);
        disp.assign(M, 0.0);
        m2.assign(M, 0.0);
        energy.assign(M, 0.0);
        norm_evo.assign((size_t)M * dnt, 0.0);
        mean_evo = disp_evo = m2_evo = energy_evo = norm_evo;
        const bool need_adr_single = dump_evo_sep == 1 && dump_adr_sep == 1;
        if (need_adr_single) adr.assign((size_t)M * d, 0.0);
        adr_avg.assign(d, 0.0);
        nj.assign(M, 0);

        if (ro_num > 0) {
            ro_evo.assign(ro_num, std::vector<std::vector<double>>(M));
            ro.assign((size_t)2 * M * ro_num, 0.0);
        }
        auto chars = [&](int is_start, int dump_id) -> int {
            int r = qjx_chars(ctx, is_start, norm.data(), mean.data(), disp.data(), m2.data(), energy.data(),
                              adr_avg.data(), need_adr_single ? adr.data() : nullptr);
            if (r < 0) return r;
            // calc_random_obs (core.cpp:694, 724) and its evo_chars_std record (core.cpp:811-815)
            if (ro_num > 0) {
                r = qjx_random_obs(ctx, ro_num, ro_seed, ro_mns, ro_type, ro.data());
                if (r < 0) return r;
                for (int tr = 0; tr < M; tr++)
                    for (int o = 0; o < ro_num; o++) {
                        const size_t k = 2 * ((size_t)tr * ro_num + o);
                        ro_evo[o][tr].push_back(ro[k]);
                        ro_evo[o][tr].push_back(ro[k + 1]);
                    }
            }
            if (is_start) mean_start = mean;
            for (int tr = 0; tr < M; tr++) {
                const size_t i = (size_t)tr * dnt + dump_id;   // evo_chars_std (core.cpp:795-816)
                norm_evo[i] = norm[tr];
                mean_evo[i] = mean[tr];
                disp_evo[i] = disp[tr];
                m2_evo[i] = m2[tr];
                energy_evo[i] = energy[tr];
            }
            return 0;
        };
        auto adr_single = [&](bool append) -> bool {
            if (!need_adr_single) return true;
            for (int tr = 0; tr < M; tr++)
                if (!save_doubles(path + "adr_" + std::to_string(tr) + sfx, &adr[(size_t)tr * d], d, append))
                    return false;
            return true;
        };
        auto adr_avg_dump = [&](bool append) -> bool {
            if (dump_adr_avg != 1) return true;
            return save_doubles(path + "adr_avg" + sfx, adr_avg.data(), d, append);
        };
        auto phi_evo = [&](bool append) -> int {
            if (dump_phi_evo != 1) return 1;
            psi.assign((size_t)M * d * 2, 0.0);
            if (qjx_get_state(ctx, psi.data(), nullptr, nullptr, nullptr, nullptr) < 0) return -1;
            for (int tr = 0; tr < M; tr++)
                if (!save_complex(path + "phi_evo_" + std::to_string(tr) + sfx, &psi[(size_t)tr * d * 2], d, append))
                    return 0;
            return 1;
        };

        // trans_process (experiment.cpp:411-446)
        API_CHECK(qjx_trans(ctx, c.num_tp_periods));
        API_CHECK(chars(1, 0));
        IO_CHECK(adr_single(false));
        if (dump_evo_avg == 1) IO_CHECK(adr_avg_dump(false));

        if (deep) {
            // StdDeepExperimentBehaviour::obser_process (experiment.cpp:806-862) with
            // ex_period_obs_deep / rk_period_obs_deep (core.cpp:114-153, 490-530): chars
            // after every sub-period, the adr average appended once per period, no phi
            // dumps, no jumps_counts stop
            const int sub = 2 * deep_steps;
            for (int pid = 0; pid < c.num_obs_periods; pid++) {
                if (c.is_pp == 1) std::cout << "dump_id: " << pid + 1 << std::endl;
                for (int k = 0; k < sub; k++) {
                    API_CHECK(qjx_run_steps(ctx, sub_steps, 1));
                    API_CHECK(chars(0, pid * sub + k + 1));
                    IO_CHECK(adr_single(true));
                }
                if (dump_evo_avg == 1) IO_CHECK(adr_avg_dump(true));
            }
        }

        // obser_process (experiment.cpp:448-542)
        if (!deep) {
            {
                int pr = phi_evo(false);
                if (pr < 0) CFG_FAIL(QJX_ERR_HIP, qjx_last_error());
                IO_CHECK(pr);
            }
            bool all_finished = false;
            for (int dump_id = 1; dump_id < dnt; dump_id++) {
                if (c.is_pp == 1) std::cout << "dump_id: " << dump_id << std::endl;
                const int n = dps[dump_id] - dps[dump_id - 1];
                const bool per_period = jump > 0 || jumps_lim > 0;
                if (per_period) {
                    for (int k = 0; k < n; k++) {
                        API_CHECK(qjx_run_periods(ctx, 1, 1));
                        if (jumps_lim > 0) {
                            API_CHECK(qjx_get_state(ctx, nullptr, nullptr, nullptr, nj.data(), nullptr));
                            all_finished = std::all_of(nj.begin(), nj.end(), [&](int v) { return v >= jumps_lim; });
                            if (all_finished) break;
                        }
                    }
                } else if (n > 0) {
                    API_CHECK(qjx_run_periods(ctx, n, 1));
                }
                API_CHECK(chars(0, dump_id));
                IO_CHECK(adr_single(true));
                {
                    int pr = phi_evo(true);
                    if (pr < 0) CFG_FAIL(QJX_ERR_HIP, qjx_last_error());
                    IO_CHECK(pr);
                }
                if (dump_evo_avg == 1) IO_CHECK(adr_avg_dump(true));
                if (all_finished && jumps_lim > 0) break;
            }
        }

        // dump_std (core.cpp:918-957)
        if (dump_obs == 1) {
            IO_CHECK(save_doubles(path + "norm" + sfx, norm.data(), M, false));
            IO_CHECK(save_doubles(path + "mean_start" + sfx, mean_start.data(), M, false));
            IO_CHECK(save_doubles(path + "mean" + sfx, mean.data(), M, false));
            IO_CHECK(save_doubles(path + "dispersion" + sfx, disp.data(), M, false));
            IO_CHECK(save_doubles(path + "m2" + sfx, m2.data(), M, false));
            IO_CHECK(save_doubles(path + "energy" + sfx, energy.data(), M, false));
        }
        // dump_phi (core.cpp:12567-12589); StdDeep writes no phi files
        if (dump_phi == 1 && !deep) {
            psi.assign((size_t)M * d * 2, 0.0);
            API_CHECK(qjx_get_state(ctx, psi.data(), nullptr, nullptr, nullptr, nullptr));
            for (int tr = 0; tr < M; tr++)
                IO_CHECK(save_complex(path + "phi_" + std::to_string(tr) + sfx, &psi[(size_t)tr * d * 2], d, false));
        }
        if (!deep) {
            int pr = phi_evo(true);
            if (pr < 0) CFG_FAIL(QJX_ERR_HIP, qjx_last_error());
            IO_CHECK(pr);
        }
        // dump_std_evo (core.cpp:1008-1084)
        if (dump_evo_sep == 1 && dump_obs == 1) {
            IO_CHECK(save_ints(path + "periods" + sfx, dps.data(), dnt, false));
            IO_CHECK(save_2d_inv(path + "norm_evo" + sfx, norm_evo, dnt, M));
            IO_CHECK(save_2d_inv(path + "mean_evo" + sfx, mean_evo, dnt, M));
            IO_CHECK(save_2d_inv(path + "dispersion_evo" + sfx, disp_evo, dnt, M));
            IO_CHECK(save_2d_inv(path + "m2_evo" + sfx, m2_evo, dnt, M));
            IO_CHECK(save_2d_inv(path + "energy_evo" + sfx, energy_evo, dnt, M));
            for (int o = 0; o < ro_num; o++)
                IO_CHECK(save_2d_complex(path + "random_obs_evo_" + std::to_string(ro_seed + o) + sfx, ro_evo[o]));
            if (jump > 0) {
                for (int tr = 0; tr < M; tr++) {
                    const int n = qjx_get_jumps(ctx, tr, nullptr, nullptr, nullptr, nullptr, 0);
                    if (n < 0) CFG_FAIL(n, qjx_last_error());
                    std::vector<double> t(n), nr(n), et(n);
                    std::vector<int> ty(n);
                    API_CHECK(qjx_get_jumps(ctx, tr, t.data(), ty.data(), nr.data(), et.data(), n));
                    IO_CHECK(save_doubles(path + "jump_times_" + std::to_string(tr) + sfx, t.data(), n, false));
                    IO_CHECK(save_ints(path + "diss_types_" + std::to_string(tr) + sfx, ty.data(), n, false));
                    if (jump > 1) {
                        IO_CHECK(save_doubles(path + "jump_norms_" + std::to_string(tr) + sfx, nr.data(), n, false));
                        IO_CHECK(save_doubles(path + "jump_etas_" + std::to_string(tr) + sfx, et.data(), n, false));
                    }
                }
            }
        }
        if (dump_evo_avg == 0) IO_CHECK(adr_avg_dump(false));
    }
done:
    if (ctx) qjx_ctx_destroy(ctx);
    if (rc != QJX_OK) qjx::set_last_error(g_cfg_err);
    if (rc == QJX_OK) {
        const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::cout << "Elapsed time: " << el << std::endl;
    }
    return rc;
}
