"""Benchmark: quantum-jump trajectory-periods/s of the dimer engine on one MI355X.

Headline workload (the precision-form configuration): Bose-Hubbard dimer N = 1000
(Hilbert dimension d = 1001), J = 1, U = 0.5, E0 = 1, A = 1.5, omega = 1,
gamma = 0.1, 1e5 trajectories, S = 200 top-level steps per period, ex_deep = 16,
start state e_0, seed 0.  A bench "step" is one top-level step T/S of the whole
ensemble, including every bisection refinement and jump it triggers; K steps are
K/S trajectory-periods per trajectory.  States and propagators are resident in
HBM before the timed region.

headline   : one transient period (FP32, from e_0) gives the common starting state;
             each form (FP32 = `value`, 3 x BF16, FP64) then takes 2 warmup steps
             and one full period (200 steps, both half periods) timed on the
             captured iteration graph, as the product runs.  Per-kernel figures
             come from a separate pass with HIP events on every launch (20 FP32
             steps: the roofline; 4 steps for the other forms).
roofline   : the FP32 step GEMM (dominant kernel); algorithmic flops = the matrix-core
             flops it executes: 3 products x 2 x 16 x 4 per (column, 16-row
             fragment, k-step of 4 complex columns) inside the fragments' band
             windows (Gauss's three-product complex multiply; the rest of the dense
             8 d^2 per application are exact zeros that are skipped) per launch of
             the event pass, over its average launch duration.  hbm_gbps / hbm_frac:
             the committed FETCH+WRITE bytes per column over the same launches.
cpu_baseline: the FP64 CPU oracle (OpenMP over trajectories, one dense zgemv per
             step as the reference) on a bounded sample started from the same
             post-transient states, with the same propagator tree.
observables_vs_fp64: each reduced form against FP64 on identical streams over 5
             periods from e_0 (1e4 trajectories).
c1, c2     : the plain FP64 drop-in at d = 101 (1000 trajectories x 100 periods,
             S = 200) and d = 501 (1e4 trajectories on the reference's S = 4 tree),
             each with its CPU oracle sample.
batched_uneven: 32 dimer instances of different dimension (seeded manifest) in
             one batch: one grouped GEMM per iteration over all of them, HIP-graph
             replay, launch shape from the tuner; also timed instance by instance.
lindblad   : the GPU density-matrix (Lindblad, rho-space RK4) integrator at N = 100
             with 1e4 RK4 steps per period, against the CPU restatement.
rk, lpn    : the Runge-Kutta trajectory path and the LPN task, each with a CPU sample.
serving    : the specified Poisson stream of ensemble jobs through the pool
             (continuous batching, admission by expected work and HBM footprint),
             p50 / p99 job latency.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "trajectory-periods/s at dim 1001 (FP32-MFMA vs FP64) + p99 ensemble-job latency"
PEAK = {"fp32": 157.3, "fp64": 78.6}   # MI355X dense matrix TFLOP/s


def make_engine(qjx, prec, N, S, D, M, **model):
    p = dict(prm_E=1.0, prm_U=0.5, prm_J=1.0, drv_ampl=1.5, drv_freq=1.0, diss_gamma=0.1)
    p.update(model)
    e = qjx.Engine(N, ex_deep=D, steps_per_period=S, precision=prec, **p)
    e.ensemble_init(M, seed=0, mns=1000000, start_type=0, start_state=0, jump=0)
    return e


def _stats_delta(s0, s1):
    return {k: s1[k] - s0[k] for k in ("matvecs", "jumps", "gemm_ms", "decide_ms", "gemm_launches",
                                       "gemm_band_flops", "gemm_flops", "iterations")}


def headline(qjx, N, M, S, D, trans_periods, warmup, steps, prof_steps, roofline_steps, forms, cpu_budget_s,
             tune_steps=60):
    """The precision-form workload, every form timed the same way.

    An FP32 engine runs `trans_periods` transient periods from e_0 (the reference's
    trans_process: no logging, then t := 0).  Its states, etas and RNG states are the
    common starting point of every form, so all of them time the same regime: each
    form, on an engine of its own started from that state, takes `warmup` steps,
    then `steps` top-level steps are timed on the captured
    iteration graph (as the product runs), then `prof_steps` more steps run with HIP
    events on every launch for the per-kernel figures (`roofline_steps` for FP32,
    whose launches give the roofline).  Observables are taken after the timed steps,
    on identical streams, so the reduced forms' errors against FP64 are correlated
    differences, not independent samples.

    No tuner trial runs in a timed region: the launch-shape tuner measures past 512
    and 4096 iterations of an engine and caches its pick per form and table shape.
    The transient run tunes the FP32 form; for the other forms a throwaway engine
    runs `tune_steps` top-level steps from the same state first, so that the timed
    engine starts from the cached shape."""
    out = {}
    e32 = make_engine(qjx, qjx.FP32, N, S, D, M)
    t0 = time.perf_counter()
    e32.trans(trans_periods)
    out["transient_s"] = time.perf_counter() - t0
    snap = e32.get_state()
    snap.pop("njumps")
    e32.close()
    names = {qjx.FP32: "fp32", qjx.BF16X3: "bf16x3", qjx.FP64: "fp64"}
    for prec in (qjx.FP32, qjx.BF16X3, qjx.FP64):
        if names[prec] not in forms:
            continue
        if prec != qjx.FP32 and tune_steps:
            et = make_engine(qjx, prec, N, S, D, M)
            et.trans(0)
            et.set_state(**snap)
            et.run_steps(tune_steps, 0)     # past the tuner's two measurements
            et.close()
        e = make_engine(qjx, prec, N, S, D, M)
        e.trans(0)
        e.set_state(**snap)
        if warmup:
            e.run_steps(warmup, 0)
        s0 = e.stats()
        t0 = time.perf_counter()
        e.run_steps(steps, 0)          # returns after the engine stream is synchronized
        dt = time.perf_counter() - t0
        s1 = e.stats()
        obs = observables(e, N, M)
        st = e.get_state(psi=False)
        e.set_profiling(True)          # resets the kernel timers
        p0 = e.stats()
        e.run_steps(roofline_steps if prec == qjx.FP32 else prof_steps, 0)
        p1 = e.stats()
        e.set_profiling(False)
        tim, prof = _stats_delta(s0, s1), _stats_delta(p0, p1)
        res = {
            "seconds": dt, "steps": steps, "warmup": warmup,
            "tp_per_s": M * steps / S / dt,
            "timing": "graph replay over %d top-level steps (%.2f periods) after %d transient period(s) and %d "
                      "warmup steps" % (steps, steps / S, trans_periods, warmup),
            "matvecs_per_traj_step": tim["matvecs"] / (M * steps),
            "jumps_per_traj_period": tim["jumps"] / (M * steps / S),
            "iterations": tim["iterations"],
            "dense_equivalent_tflops": tim["gemm_flops"] / dt / 1e12,
            "profiled_steps": prof,
            "gemm_band_tflops": prof["gemm_band_flops"] / (prof["gemm_ms"] * 1e-3) / 1e12 if prof["gemm_ms"] else None,
            "gemm_share_of_profiled": None,
            "gemm_bn": p1["gemm_bn"], "gemm_grid": p1["gemm_grid"],
            "gemm_kernel": gemm_kernel_name(prec, p1["gemm_bn"], p1["gemm_waves"]),
            "observables": obs,
            "rng": st["rng"],
            # the timed engine's whole life (tuner trials if it tuned, warmup, timed and
            # profiled steps): what a kernel trace or counter pass of the run sees
            # with no transient period
            "engine_totals": {"matvecs": p1["matvecs"], "gemm_band_flops": p1["gemm_band_flops"],
                              "iterations": p1["iterations"]},
        }
        if prec == qjx.FP64 and cpu_budget_s > 0:
            out["cpu_baseline"] = cpu_baseline(e, N, S, D, cpu_budget_s, start=snap)
        e.close()
        out[names[prec]] = res
    del snap
    for k in ("fp32", "bf16x3"):
        if k in out and "fp64" in out:
            out[k]["same_jump_count_as_fp64"] = float((out[k]["rng"] == out["fp64"]["rng"]).mean())
    for k in ("fp32", "bf16x3", "fp64"):
        if k in out:
            out[k].pop("rng")
    return out


def observables_leg(qjx, N, M, S, D, periods, forms):
    """<n1 - n2>/N and purity of each reduced form against the FP64 run after every
    period up to `periods`, from e_0 on identical streams (graph replay), and the
    share of trajectories whose jump count still equals the FP64 one (equal final
    RNG states).  Once a trajectory's crossing test falls the other way in one form
    its later jumps differ, and the forms' ensembles become independent samples:
    then the error is compared with the combined standard error."""
    names = {qjx.FP32: "fp32", qjx.BF16X3: "bf16x3", qjx.FP64: "fp64"}
    per, rng, secs = {}, {}, {}
    for prec in (qjx.FP64, qjx.FP32, qjx.BF16X3):
        if names[prec] not in forms and prec != qjx.FP64:
            continue
        e = make_engine(qjx, prec, N, S, D, M)
        e.trans(0)
        t0 = time.perf_counter()
        rows = []
        for _ in range(periods):
            e.run_periods(1, 0)
            rows.append(observables(e, N, M))
        secs[names[prec]] = time.perf_counter() - t0
        rng[names[prec]] = e.get_state(psi=False)["rng"]
        per[names[prec]] = rows
        e.close()
    out = {"trajectories": M, "periods": periods, "seconds": secs, "fp64": per["fp64"]}
    for k in ("fp32", "bf16x3"):
        if k not in per:
            continue
        errs = []
        for a, b in zip(per[k], per["fp64"]):
            d = a["n1_minus_n2_over_N"] - b["n1_minus_n2_over_N"]
            sg = (a["sigma"] ** 2 + b["sigma"] ** 2) ** 0.5
            errs.append({"n1_minus_n2_over_N_error": d, "combined_sigma": sg, "error_in_sigma": d / sg if sg else None,
                         "purity_error": a["purity"] - b["purity"]})
        out[k] = {"by_period": errs, "same_jump_count_as_fp64": float((rng[k] == rng["fp64"]).mean()),
                  "max_abs_n1_minus_n2_over_N_error": max(abs(x["n1_minus_n2_over_N_error"]) for x in errs),
                  "max_abs_purity_error": max(abs(x["purity_error"]) for x in errs)}
    return out


def drop_in_leg(qjx, N, M, S, D, trans_periods, periods, cpu_budget_s, cpu_traj, label, cpu_periods=1,
                physics_periods=0):
    """The plain FP64 drop-in on one configuration: M trajectories of the dimer at N
    on the tree (S, D), `trans_periods` transient periods, then `periods` timed
    periods on graph replay; the CPU oracle (OpenMP over trajectories, one zgemv per
    step as the reference, the same tree) on a bounded sample from the start state."""
    e = make_engine(qjx, qjx.FP64, N, S, D, M)
    t0 = time.perf_counter()
    e.trans(trans_periods)
    trans_s = time.perf_counter() - t0
    start = None
    if cpu_budget_s > 0 and trans_periods > 0:
        start = e.get_state()
    s0 = e.stats()
    physics = None
    if physics_periods:
        # the ensemble against its Lindblad oracle (RK4, 1e4 steps per period, CQdiss's
        # per-stage drive) after physics_periods of the timed periods (untimed check)
        t0 = time.perf_counter()
        e.run_periods(physics_periods, 0)
        dt = time.perf_counter() - t0
        lb = qjx.Lindblad(N, steps_per_period=10000, diss_gamma=0.1, per_stage=1, prm_E=1.0, prm_U=0.5, prm_J=1.0,
                          drv_ampl=1.5, drv_freq=1.0)
        lb.run_periods(trans_periods + physics_periods)
        ob = lb.observables()
        lb.close()
        c = e.chars()
        physics = {"periods": trans_periods + physics_periods, "trajectories": M}
        for k in ("mean", "energy"):
            v, sig = float(c[k].mean()), float(c[k].std() / M ** 0.5)
            physics[k] = {"ensemble": v, "lindblad": ob[k], "sigma": sig, "error_in_sigma": (v - ob[k]) / sig}
        t0 = time.perf_counter()
        e.run_periods(periods - physics_periods, 0)
        dt += time.perf_counter() - t0
    else:
        t0 = time.perf_counter()
        e.run_periods(periods, 0)
        dt = time.perf_counter() - t0
    s1 = e.stats()
    d = _stats_delta(s0, s1)
    res = {"config": label, "N": N, "dim": N + 1, "trajectories": M, "steps_per_period": S, "ex_deep": D,
           "transient_periods": trans_periods, "timed_periods": periods, "seconds": dt, "transient_s": trans_s,
           "tp_per_s": M * periods / dt, "matvecs_per_traj_period": d["matvecs"] / (M * periods),
           "jumps_per_traj_period": d["jumps"] / (M * periods), "iterations": d["iterations"],
           "dense_equivalent_tflops": d["gemm_flops"] / dt / 1e12, "lindblad_check": physics}
    if cpu_budget_s > 0:
        res["cpu_baseline"] = cpu_baseline(e, N, S, D, cpu_budget_s, ntraj=cpu_traj, start=start,
                                           max_steps=cpu_periods * S)
        res["gpu_over_cpu"] = res["tp_per_s"] / res["cpu_baseline"]["value"]
    e.close()
    return res


def gemm_kernel_name(prec, bn, waves):
    """Name of the step GEMM kernel the launch shape selects (as a kernel trace shows it)."""
    if prec == 0:
        return "qjx::gemm_f64<%d>" % bn
    # complex-pair GEMM forms: 1 = 3xBF16, 2 = FP32 (16-row fragments)
    return "qjx::gemm_cp<%d, %d, %d>" % (1 if prec == 2 else 2, bn, 4 if (bn == 256 or waves == 8) else 2)


def observables(e, N, M):
    """<n1 - n2>/N over the ensemble (k = n2 in the dimer basis, so n1 - n2 = N - 2k)
    with its standard error, and the purity of the ensemble density matrix."""
    import numpy as np
    c = e.chars()
    x = (N - 2.0 * c["mean"]) / N
    _, pur = e.density(rho=False)
    return {"n1_minus_n2_over_N": float(x.mean()), "sigma": float(x.std() / np.sqrt(M)), "purity": pur}


def observable_errors(form, ref):
    """Error of a reduced-precision run's observables against the FP64 run (same
    workload, seeds and steps; independent jump sequences, so the mean's error is
    compared with the combined standard error)."""
    if not form or not ref:
        return None
    a, b = form["observables"], ref["observables"]
    d = a["n1_minus_n2_over_N"] - b["n1_minus_n2_over_N"]
    s = (a["sigma"] ** 2 + b["sigma"] ** 2) ** 0.5
    return {"n1_minus_n2_over_N_error": d, "combined_sigma": s, "error_in_sigma": d / s if s > 0 else None,
            "purity_error": a["purity"] - b["purity"]}


TRAFFIC_PROFILE = os.path.join("profiles", "step_gemm_traffic_dimer_d1001_M1e5.md")


def measured_traffic_per_column(form="fp32"):
    """HBM bytes per propagator application of a form's step GEMM, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated against a copy kernel)."""
    try:
        vals = {}
        for line in open(os.path.join(REPO, TRAFFIC_PROFILE)):
            if line.startswith(form + "_traffic_bytes_per_column_"):
                k, v = line.split(":")
                vals[k.strip()[len(form) + 1:]] = float(v)
        return vals["traffic_bytes_per_column_fetch"] + vals["traffic_bytes_per_column_write"]
    except (OSError, KeyError, ValueError):
        return None


def cpu_baseline(engine, N, S, D, budget_s, ntraj=0, start=None, max_steps=0):
    """FP64 oracle on the host cores, bounded sample, same propagator tree: ntraj
    trajectories (default 2 per thread) from the start state, top-level steps until
    the budget is spent or a period is done."""
    import numpy as np
    from oracle import oracle as orc
    threads = orc.default_threads()
    m = orc.model(N, prm_E=1.0, prm_U=0.5, prm_J=1.0, drv_ampl=1.5, drv_freq=1.0, gamma=0.1, S=S, D=D)
    ntraj = ntraj or 2 * threads
    o = orc.Ensemble(m, ntraj, seed=0, jump=0, threads=threads, build_propagators=False)
    for b in range(2):
        for lvl in range(D):
            o.set_propagator(b, lvl, engine.get_propagator(b, lvl))
    o.trans(0)
    where = "the start state"
    if start is not None:
        o.set_state(psi=start["psi"][:ntraj], eta=start["eta"][:ntraj], t=np.zeros(ntraj),
                    rng=start["rng"][:ntraj])
        where = "the GPU run's post-transient states"
    t0 = time.perf_counter()
    nsteps = 0
    while True:
        o.steps(1, 0)
        nsteps += 1
        if time.perf_counter() - t0 >= budget_s or nsteps >= (max_steps or S):
            break
    dt = time.perf_counter() - t0
    return {"value": ntraj * nsteps / S / dt, "unit": "trajectory-periods/s", "cores": threads,
            "kind": "port",
            "sample": "%d trajectories x %d top-level steps (%d per period) from %s, d=%d, FP64 oracle "
                      "(OpenMP over trajectories, one dense zgemv per step), %.1f s"
                      % (ntraj, nsteps, S, where, N + 1, dt),
            "matvecs_per_traj_step": o.num_matvecs() / (ntraj * nsteps)}


C4_SEED = 12345


def batched_instances(n=32, seed=C4_SEED):
    """The batched uneven-ensembles workload: n dimer instances, N uniform on
    {50..2000}, A ~ U[0.5, 3.5], U ~ U[0.1, 1.0], gamma ~ U[0.01, 0.5],
    trajectories log-uniform on [1e3, 5e4]; numpy PCG64 seeded with `seed`
    (the manifest is printed with the results)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        N = int(rng.integers(50, 2001))
        A = float(rng.uniform(0.5, 3.5))
        U = float(rng.uniform(0.1, 1.0))
        g = float(rng.uniform(0.01, 0.5))
        M = int(round(10 ** rng.uniform(3.0, np.log10(5e4))))
        out.append({"N": N, "M": M, "drv_ampl": A, "prm_U": U, "diss_gamma": g, "seed": i})
    return out


def _instance_engine(qjx, p, S, D, prec):
    e = qjx.Engine(p["N"], prm_E=1.0, prm_U=p["prm_U"], prm_J=1.0, drv_ampl=p["drv_ampl"], drv_freq=1.0,
                   diss_gamma=p["diss_gamma"], ex_deep=D, steps_per_period=S, precision=prec)
    e.ensemble_init(p["M"], seed=p["seed"], mns=1000000, start_type=0, start_state=0, jump=0)
    e.trans(0)
    return e


def batched(qjx, S, D, warmup, steps, prec, alone):
    """All instances in one batch: every iteration is one persistent grouped GEMM
    over the 32 ensembles (HIP-graph replay, launch shape from the tuner).  With
    `alone`, the same steps are also run instance by instance for comparison."""
    inst = batched_instances()
    t0 = time.perf_counter()
    engines = [_instance_engine(qjx, p, S, D, prec) for p in inst]
    build_s = time.perf_counter() - t0
    b = qjx.Batch(engines)
    if warmup:
        b.run_steps(warmup, 0)
    s0 = b.stats()
    t0 = time.perf_counter()
    b.run_steps(steps, 0)
    dt = time.perf_counter() - t0
    s1 = b.stats()
    tp = sum(p["M"] for p in inst) * steps / S
    res = {"instances": len(inst), "manifest_seed": C4_SEED, "trajectories": sum(p["M"] for p in inst),
           "dims": [p["N"] + 1 for p in inst], "steps": steps, "warmup": warmup, "seconds": dt,
           "tp_per_s": tp / dt, "matvecs": s1["matvecs"] - s0["matvecs"],
           "effective_band_tflops": (s1["gemm_band_flops"] - s0["gemm_band_flops"]) / dt / 1e12,
           "effective_dense_equivalent_tflops": (s1["gemm_flops"] - s0["gemm_flops"]) / dt / 1e12,
           "iterations": s1["iterations"] - s0["iterations"], "gemm_bn": s1["gemm_bn"],
           "gemm_grid": s1["gemm_grid"], "gemm_kernel": gemm_kernel_name(prec, s1["gemm_bn"], s1["gemm_waves"]),
           "build_s": build_s, "manifest": inst}
    b.close()
    if alone:
        t_alone = 0.0
        for e in engines:
            if warmup:
                e.run_steps(warmup, 0)
            t0 = time.perf_counter()
            e.run_steps(steps, 0)
            t_alone += time.perf_counter() - t0
        res["alone_seconds"] = t_alone
        res["alone_tp_per_s"] = tp / t_alone
        res["batch_speedup"] = t_alone / dt
    for e in engines:
        e.close()
    return res


def lindblad(qjx, N=100, steps=10000, periods=2, cpu_steps=200):
    """The Lindblad density-matrix integrator (rho-space RK4, FP64) on the
    trajectory workload's model at N = 100 with 1e4 RK4 steps per period (the
    configuration the reference's f-basis solver is run on for this model), on the
    GPU and, for a bounded number of steps, on the host with the CPU restatement
    (single thread).  Both produce the same bits."""
    import numpy as np
    from oracle import oracle as orc
    params = dict(prm_E=1.0, prm_U=0.5, prm_J=1.0, drv_ampl=1.5, drv_freq=1.0)
    lb = qjx.Lindblad(N, steps_per_period=steps, diss_gamma=0.1, per_stage=1, **params)
    lb.run_periods(1)                       # warm-up: graph capture, caches
    t0 = time.perf_counter()
    lb.run_periods(periods)
    dt = time.perf_counter() - t0
    ob = lb.observables()
    lb.close()
    # the CPU sample: cpu_steps RK4 steps of the same size and cost, as one period of
    # a drive whose frequency is scaled by steps / cpu_steps
    cpu_params = dict(params, drv_freq=params["drv_freq"] * steps / cpu_steps)
    m = orc.model(N, gamma=0.1, **cpu_params)
    rho = np.zeros((N + 1, N + 1), complex)
    rho[0, 0] = 1.0
    t0 = time.perf_counter()
    orc.lindblad_evolve_drive(m, rho, 1, steps_per_period=cpu_steps, per_stage=1)
    dtc = time.perf_counter() - t0
    cpu_steps_per_s = cpu_steps / dtc
    gpu_steps_per_s = periods * steps / dt
    return {"N": N, "dim": N + 1, "rk4_steps_per_period": steps, "gpu_periods_per_s": periods / dt,
            "gpu_rk4_steps_per_s": gpu_steps_per_s, "cpu_rk4_steps_per_s": cpu_steps_per_s,
            "cpu_sample": "%d RK4 steps at d=%d, one host thread, the CPU restatement" % (cpu_steps, N + 1),
            "gpu_over_cpu": gpu_steps_per_s / cpu_steps_per_s,
            "observables_after": {"mean": ob["mean"], "energy": ob["energy"], "purity": ob["purity"]}}


RK_FLOPS_PER_ENTRY_STEP = 182.0   # FP64 flops per state entry per RK step (fma = 2)


def rk_leg(qjx, N, rk_ns, M, steps, cpu_budget_s, drv_type=0, with_cpu=True):
    """The Runge-Kutta trajectory path (prop_id 1): M trajectories of the dimer at
    N with rk_ns RK4 steps per period, `steps` steps timed on the GPU (kernel time
    from HIP events on the engine stream and wall time), and the CPU restatement on
    the host cores for a bounded sample of the same steps (same bits)."""
    from oracle import oracle as orc
    params = dict(prm_E=1.0, prm_U=0.5, prm_J=1.0, drv_ampl=1.5, drv_freq=1.0)
    e = qjx.Engine(N, diss_gamma=0.1, prop="rk", rk_ns=rk_ns, drv_type=drv_type, **params)
    e.ensemble_init(M, seed=0, jump=0)
    e.trans(0)
    e.run_steps(min(steps, 200), 0)              # warm-up (code objects, caches)
    s0 = e.stats()
    e.set_profiling(1)
    t0 = time.perf_counter()
    e.run_steps(steps, 0)
    wall = time.perf_counter() - t0
    s1 = e.stats()
    e.close()
    kernel_s = s1["gemm_ms"] * 1e-3
    rhs = s1["matvecs"] - s0["matvecs"]
    flops = RK_FLOPS_PER_ENTRY_STEP * (N + 1) * rhs / 4.0
    res = {"N": N, "dim": N + 1, "rk_ns": rk_ns, "drv_type": drv_type, "trajectories": M, "steps": steps,
           "trajectory_periods_per_s": M * steps / rk_ns / wall,
           "trajectory_steps_per_s": M * steps / wall,
           "wall_s": wall, "kernel_s": kernel_s, "launches": s1["gemm_launches"],
           "integrations_per_step": rhs / 4.0 / (M * steps), "jumps": s1["jumps"] - s0["jumps"],
           "roofline": {"bound": "fp64-valu", "achieved": flops / kernel_s / 1e12 if kernel_s > 0 else None,
                        "peak": PEAK["fp64"], "unit": "TFLOP/s",
                        "frac": flops / kernel_s / 1e12 / PEAK["fp64"] if kernel_s > 0 else None,
                        "flops_note": "%.0f FP64 flops per state entry per RK4 integration (4 stencil stages "
                                      "of the pentadiagonal H_part + diagonal drive, RK updates, norm)"
                                      % RK_FLOPS_PER_ENTRY_STEP,
                        "kernel": "rk_kernel (32 or 64 lanes per trajectory, state in registers)"}}
    if with_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1
        threads = min(threads, os.cpu_count() or threads)
        m = orc.model(N, gamma=0.1, D=1, **params)
        ntraj = 2 * threads
        o = orc.Ensemble(m, ntraj, seed=0, jump=0, threads=threads, build_propagators=False)
        o.rk_setup(rk_ns, drv_type, 0.0)
        o.trans(0)
        t0 = time.perf_counter()
        n = 0
        chunk = max(1, rk_ns // 100)
        while time.perf_counter() - t0 < cpu_budget_s and n < 100 * rk_ns:
            o.steps(chunk, 0)
            n += chunk
        dtc = time.perf_counter() - t0
        cpu = ntraj * n / rk_ns / dtc
        res["cpu_baseline"] = {"value": cpu, "unit": "trajectory-periods/s", "cores": threads, "kind": "port",
                               "sample": "%d trajectories x %d RK steps at d=%d, CPU restatement, %.1f s"
                                         % (ntraj, n, N + 1, dtc)}
        res["gpu_over_cpu"] = res["trajectory_periods_per_s"] / cpu
    return res


def lpn_leg(qjx, M, periods, cpu_periods, with_cpu=True):
    """The LPN task (task_id 7) on the reference's own LPN parameter set
    (source/cpp/QJX/QJX/params.txt and config.txt: N = 10, E = 0, U = 0.3, J = 1, sine
    drive A = 3.4, omega = 1, gamma = 0.1, lpn_type -1, lambda_per_periods 1, RK with
    rk_ns = 10000) with M trajectories (M / 2 pairs): GPU propagation plus host
    bookkeeping per period, against the CPU restatement on the host cores."""
    from oracle import oracle as orc
    params = dict(prm_E=0.0, prm_U=0.3, prm_J=1.0, drv_ampl=3.4, drv_freq=1.0)
    lpn = dict(lpn_type=-1, lambda_per_periods=1)
    e = qjx.Engine(10, diss_gamma=0.1, prop="rk", rk_ns=10000, drv_type=1, **params)
    e.ensemble_init(M, seed=1, jump=0)
    L = qjx.Lpn(e, **lpn)
    L.trans(0)
    L.run_periods(1)                                    # warm-up
    t0 = time.perf_counter()
    L.run_periods(periods)
    wall = time.perf_counter() - t0
    lam = L.get()["lambda_now"][M // 2:]
    L.close()
    e.close()
    res = {"N": 10, "dim": 11, "trajectories": M, "pairs": M // 2, "rk_ns": 10000, "periods": periods,
           "lpn_periods_per_s": periods / wall, "trajectory_periods_per_s": M * periods / wall,
           "lambda_mean": float(lam.mean()), "lambda_std": float(lam.std())}
    if with_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1
        threads = min(threads, os.cpu_count() or threads)
        m = orc.model(10, gamma=0.1, D=1, **params)
        o = orc.Ensemble(m, M, seed=1, jump=0, threads=threads, build_propagators=False)
        o.rk_setup(10000, 1, 0.0)
        o.lpn_setup(**lpn)
        o.lpn_trans(0)
        t0 = time.perf_counter()
        for _ in range(cpu_periods):
            o.lpn_period()
        dtc = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": M * cpu_periods / dtc, "unit": "trajectory-periods/s", "cores": threads,
                               "kind": "port", "sample": "%d trajectories x %d LPN periods, CPU restatement, %.1f s"
                                                         % (M, cpu_periods, dtc)}
        res["gpu_over_cpu"] = res["trajectory_periods_per_s"] / res["cpu_baseline"]["value"]
    return res


SERVE_MIXES = {
    # the serving configuration as specified
    "specified": {"N": [100, 250, 500, 1000, 2000], "logM": (3.0, 5.0), "periods": (50, 500)},
    # same arrival process, jobs small enough for the offered load to fit the device
    "scaled": {"N": [100, 250, 500], "logM": (2.0, 3.0), "periods": (1, 5)},
}


def serve_mix(qjx, pool, seconds, rate, seed, mix, probe=None, probe_at=None, check=False):
    """Drive `pool` with a Poisson stream of ensemble jobs for `seconds` on the pool
    clock (arrivals stop then).  Each job is a dimer parameter point: N from the
    mix, trajectories log-uniform, periods uniform, A in [0.5, 3.5], U in [0.1, 1.0],
    gamma in [0.01, 0.5]; the reference's QJ tree (4 top-level steps per period,
    ex_deep 16).  `probe` (submit keywords) adds one extra job at `probe_at` s.
    Returns the jobs (dicts with the pool's id) and the largest HBM reservation seen;
    with `check`, every cycle asserts that the reservation stays within the budget."""
    import numpy as np
    spec = SERVE_MIXES[mix]
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=int(rate * seconds * 2) + 16))
    arrivals = arrivals[arrivals < seconds]
    jobs, peak = [], 0
    t0 = pool.clock()
    nxt = 0
    while True:
        now = pool.clock() - t0
        while nxt < len(arrivals) and arrivals[nxt] <= now:
            N = int(rng.choice(spec["N"]))
            M = int(round(10 ** rng.uniform(*spec["logM"])))
            P = int(rng.integers(spec["periods"][0], spec["periods"][1] + 1))
            jid = pool.submit(N, M, P, seed=int(rng.integers(0, 1 << 20)), drv_ampl=float(rng.uniform(0.5, 3.5)),
                              prm_U=float(rng.uniform(0.1, 1.0)), diss_gamma=float(rng.uniform(0.01, 0.5)),
                              steps_per_period=4, ex_deep=16)
            jobs.append({"id": jid, "N": N, "M": M, "P": P, "probe": False})
            nxt += 1
        if probe is not None and probe_at is not None and now >= probe_at:
            jobs.append({"id": pool.submit(**probe), "N": probe["N"], "M": probe["num_trajectories"],
                         "P": probe["n_periods"], "probe": True})
            probe_at = None
        if now >= seconds:
            break
        r, q = pool.step()
        u = pool.usage()
        peak = max(peak, u["used_bytes"])
        if check:
            assert u["used_bytes"] <= u["budget_bytes"], u
            assert u["failed"] == 0, u
        if r == 0 and q == 0 and nxt >= len(arrivals) and probe_at is None:
            break
    return jobs, peak


def serving(qjx, seconds, rate, seed, budget_gib, capacity, mix="specified"):
    """Poisson stream of ensemble jobs through the pool (continuous batching, FP32
    form); latency = completion - submission.  p50 / p99 are over the jobs completed
    within the window; `p99_all_lower_bound_s` also counts every job still queued or
    running at the end with its age then (a lower bound of its latency)."""
    pool = qjx.Pool(qjx.FP32, hbm_budget_gib=budget_gib, max_trajectories=capacity)
    t0 = pool.clock()
    jobs, peak = serve_mix(qjx, pool, seconds, rate, seed, mix)
    t_end = pool.clock() - t0
    done, lat, lat_all, tp, admitted = 0, [], [], 0.0, 0
    for j in jobs:
        res = pool.result(j["id"])
        if res["status"] != 0 or res["t_admit"] > 0:
            admitted += 1
        if res["status"] == 2:
            done += 1
            lat.append(res["t_done"] - res["t_submit"])
            lat_all.append(lat[-1])
            tp += j["M"] * j["P"]
        else:
            lat_all.append(t_end + t0 - res["t_submit"])
    u = pool.usage()
    pool.close()
    spec = SERVE_MIXES[mix]

    def pct(v, p):
        v = sorted(v)
        return v[min(len(v) - 1, int(p * len(v)))] if v else None
    return {"mix": mix, "jobs": {"N": spec["N"], "trajectories_log10": spec["logM"], "periods": spec["periods"]},
            "offered_jobs_per_s": rate, "window_s": t_end, "submitted": len(jobs), "admitted": admitted,
            "completed": done, "failed": u["failed"], "running_at_end": u["running"],
            "queued_at_end": u["queued"] + u["building"],
            "p50_latency_s": pct(lat, 0.5), "p99_latency_s": pct(lat, 0.99),
            "p50_all_lower_bound_s": pct(lat_all, 0.5), "p99_all_lower_bound_s": pct(lat_all, 0.99),
            "peak_reserved_gib": peak / 2 ** 30, "budget_gib": u["budget_bytes"] / 2 ** 30,
            "completed_trajectory_periods_per_s": tp / t_end if t_end > 0 else 0.0}


LEGS = ("headline", "observables", "c1", "c2", "batch", "lindblad", "rk", "lpn", "serving")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200, help="timed top-level steps per form (200 = one period)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--trans", type=int, default=1, help="transient periods before the timed steps")
    ap.add_argument("--roofline-steps", type=int, default=20,
                    help="FP32 steps with HIP events on every launch (the roofline pass)")
    ap.add_argument("--prof-steps", type=int, default=4, help="profiled steps of the other forms")
    ap.add_argument("--N", type=int, default=1000)
    ap.add_argument("--traj", type=int, default=100000)
    ap.add_argument("--S", type=int, default=200)
    ap.add_argument("--D", type=int, default=16)
    ap.add_argument("--forms", default="fp32,bf16x3,fp64")
    ap.add_argument("--legs", default=",".join(LEGS), help="comma list of " + ",".join(LEGS))
    ap.add_argument("--obs-traj", type=int, default=10000)
    ap.add_argument("--obs-periods", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=4)
    ap.add_argument("--batch-alone", type=int, default=1,
                    help="also time the batched workload instance by instance")
    ap.add_argument("--rk-cpu-seconds", type=float, default=8.0)
    ap.add_argument("--serve-seconds", type=float, default=90.0, help="serving window on the pool clock")
    ap.add_argument("--serve-mix", default="specified", choices=sorted(SERVE_MIXES))
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-budget-gib", type=float, default=0.0)
    ap.add_argument("--serve-capacity", type=int, default=2000000)
    ap.add_argument("--peak-probe", action="store_true", help="only the measured device ceilings")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (read + write), the "
                         "calibration kernel for the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.copy_probe_mib:
        import torch
        n = args.copy_probe_mib * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            torch.mul(x, 1.0, out=y)       # one elementwise kernel: n*4 bytes read, n*4 written
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4,
                          "launches": 3}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the trajectory engine runs on a single device "
              "(splitting the trajectories over several GPUs is designed, not implemented).",
              file=sys.stderr)
        sys.exit(2)

    from os_d_amd import qjx
    N, M, S, D = args.N, args.traj, args.S, args.D
    if args.peak_probe:
        print(json.dumps({"measured_peaks": qjx.peak_probe()}), flush=True)
        return
    legs = set(x for x in args.legs.split(",") if x)
    forms = set(x for x in args.forms.split(",") if x)
    cpu_s = 0.0 if args.no_cpu else args.cpu_seconds
    with_cpu = not args.no_cpu

    def say(msg):
        print("bench: " + msg, file=sys.stderr, flush=True)

    head = None
    if "headline" in legs:
        say("headline d=%d M=%d S=%d forms %s" % (N + 1, M, S, ",".join(sorted(forms))))
        head = headline(qjx, N, M, S, D, args.trans, args.warmup, args.steps, args.prof_steps,
                        args.roofline_steps, forms, cpu_s if "fp64" in forms else 0.0)
    obs = None
    if "observables" in legs:
        say("observables vs FP64: M=%d x %d periods" % (args.obs_traj, args.obs_periods))
        obs = observables_leg(qjx, N, args.obs_traj, S, D, args.obs_periods, forms)
    c1 = c2 = None
    if "c1" in legs:
        # the plain CPU configuration: N=100, 1000 trajectories x 100 periods, S=200
        say("c1 d=101")
        c1 = drop_in_leg(qjx, 100, 1000, 200, 16, 0, 100, min(cpu_s, 10.0), 0,
                         "N=100 (d=101), 1000 trajectories x 100 periods, S=200, ex_deep 16, FP64", cpu_periods=100,
                         physics_periods=10)
    if "c2" in legs:
        # the FP64 drop-in at N=500 on the reference's tree (4 top-level steps per period)
        say("c2 d=501")
        c2 = drop_in_leg(qjx, 500, 10000, 4, 16, 1, 2, min(cpu_s, 15.0), 0,
                         "N=500 (d=501), 1e4 trajectories, reference QJ tree (S=4, ex_deep 16), FP64: "
                         "1 transient + 2 timed periods of the specified 500")
    batch = None
    if "batch" in legs and args.batch_steps > 0:
        say("batched uneven ensembles")
        batch = batched(qjx, S, D, 2, args.batch_steps, qjx.FP32, bool(args.batch_alone))
    lind = None
    if "lindblad" in legs:
        say("lindblad")
        lind = lindblad(qjx)
    rk = None
    if "rk" in legs:
        say("rk")
        # the reference's CPU configuration (N=100) with its default rk_ns, and the
        # headline dimension with a step that keeps RK4 stable (h |H| ~ 0.5)
        rk = {"d101": rk_leg(qjx, 100, 10000, 10000, 20000, args.rk_cpu_seconds, with_cpu=with_cpu),
              "d1001": rk_leg(qjx, 1000, 40000, 4096, 4000, args.rk_cpu_seconds, with_cpu=with_cpu)}
    lpn = None
    if "lpn" in legs:
        say("lpn")
        # the reference's 100 trajectories, and enough pairs to fill the device
        lpn = {"M100": lpn_leg(qjx, 100, 5, 1, with_cpu), "M4096": lpn_leg(qjx, 4096, 5, 1, with_cpu)}
    serve = None
    if "serving" in legs and args.serve_seconds > 0:
        say("serving %s mix, %.0f s" % (args.serve_mix, args.serve_seconds))
        serve = serving(qjx, args.serve_seconds, args.serve_rate, 12345, args.serve_budget_gib,
                        args.serve_capacity, args.serve_mix)
    peaks = qjx.peak_probe()

    line = {
        "metric": METRIC,
        "value": None,
        "unit": "trajectory-periods/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (start state e_0, seeded MCG59 streams; no dataset)",
        "config": {"workload": "BH dimer N=%d (d=%d), %d trajectories, S=%d steps/period, ex_deep=%d, "
                               "FP32 propagation + FP64 norms/bisection/jumps; one period timed after %d "
                               "transient period(s)" % (N, N + 1, M, S, D, args.trans),
                   "trajectories": M, "dim": N + 1, "steps_per_period": S, "ex_deep": D,
                   "step": "one top-level step T/S of the whole ensemble incl. refinement"},
        "measured_peaks": peaks,
    }
    if head and "fp32" in head:
        r32 = head["fp32"]
        prof = r32["profiled_steps"]
        launches = max(1, prof["gemm_launches"])
        avg_ms = prof["gemm_ms"] / launches
        cols_per_launch = prof["matvecs"] / launches
        flops_per_launch = prof["gemm_band_flops"] / launches
        per_col = measured_traffic_per_column()
        achieved = flops_per_launch / (avg_ms * 1e-3) / 1e12 if avg_ms > 0 else 0.0
        line["value"] = r32["tp_per_s"]
        line["ms_per_step"] = 1e3 * r32["seconds"] / args.steps
        line["roofline"] = {
            "bound": "mfma", "achieved": achieved, "peak": PEAK["fp32"], "unit": "TFLOP/s",
            "frac": achieved / PEAK["fp32"],
            "traffic": per_col * cols_per_launch if (per_col and N == 1000) else None,
            "peak_measured": peaks["f32_mfma_tflops"], "frac_of_measured": achieved / peaks["f32_mfma_tflops"],
            "traffic_source": TRAFFIC_PROFILE + " (fp32 bytes/column x columns/launch)",
            "algorithmic_bytes_per_launch": 2 * 4 * 2 * 1024 * cols_per_launch if N == 1000 else None,
            "kernel": "%s (grouped step GEMM, gathered columns)" % r32["gemm_kernel"],
            "launches": launches, "avg_launch_ms": avg_ms, "flops_per_launch": flops_per_launch,
            "columns_per_launch": cols_per_launch,
            "flops_note": "3 products x 2 x 16 x 4 per (column, 16-row fragment, k-step of 4 complex columns) "
                          "inside the fragments' band windows (three-product complex multiply); launches of the "
                          "%d-step HIP-event pass after the timed period" % args.roofline_steps,
            "hbm_gbps": (per_col * cols_per_launch / (avg_ms * 1e-3) / 1e9) if (per_col and N == 1000 and avg_ms > 0) else None,
            "hbm_frac": (per_col * cols_per_launch / (avg_ms * 1e-3) / 8e12) if (per_col and N == 1000 and avg_ms > 0) else None,
            "dense_equivalent_flops_per_launch": 8.0 * (N + 1) ** 2 * cols_per_launch,
        }
        line["fp32"] = r32
    if head:
        line["transient_s"] = head.get("transient_s")
        for k in ("bf16x3", "fp64"):
            if k in head:
                line[k + "_value"] = head[k]["tp_per_s"]
                line[k] = head[k]
        for k in ("fp32", "bf16x3", "fp64"):
            per_col = measured_traffic_per_column(k)
            if k in head and per_col and N == 1000:
                # L2-miss bytes per second of the form's GEMM over the timed period
                cols_per_s = head[k]["tp_per_s"] * head[k]["matvecs_per_traj_step"] * S
                head[k]["gemm_traffic_tb_per_s"] = per_col * cols_per_s / 1e12
        if "fp32" in head and "fp64" in head:
            line["fp32_over_fp64"] = head["fp32"]["tp_per_s"] / head["fp64"]["tp_per_s"]
        if "bf16x3" in head and "fp64" in head:
            line["bf16x3_over_fp64"] = head["bf16x3"]["tp_per_s"] / head["fp64"]["tp_per_s"]
        line["cpu_baseline"] = head.get("cpu_baseline")
        if line["cpu_baseline"] and line["value"]:
            line["fp32_over_cpu"] = line["value"] / line["cpu_baseline"]["value"]
        line["observables_after_timed_period_vs_fp64"] = {
            k: observable_errors(head.get(k), head.get("fp64")) for k in ("fp32", "bf16x3")}
    line["observables_vs_fp64"] = obs
    line["c1_fp64_d101"] = c1
    line["c2_fp64_d501"] = c2
    line["batched_uneven"] = batch
    line["lindblad"] = lind
    line["rk_trajectories"] = rk
    line["lpn"] = lpn
    line["serving"] = serve
    if serve:
        line["p50_job_latency_s"] = serve["p50_latency_s"]
        line["p99_job_latency_s"] = serve["p99_latency_s"]
        line["p99_job_latency_note"] = ("%s mix, %.0f s window: over the %d jobs completed in the window; "
                                        "p99_all_lower_bound_s counts the unfinished jobs with their age"
                                        % (serve["mix"], serve["window_s"], serve["completed"]))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
