"""Benchmark of the PDOT hot path on one MI355X.

Workload: the 10000 x 10000 Gaussian-shift transport problem (source N(0,I2),
target N((2,0),I2), squared Euclidean cost / max, U(0,1) masses, seeded
synthetic data), solved by restarted adaptive PDHG.

A step is one PDHG iteration (one take_step! pass over C, X and the running
average).  The timed region runs K steps of the FP64 form (the reference's
precision) on inputs resident in HBM, with HIP events recorded on the
library's stream around every launch of the streaming step kernel, so the
same run gives the iteration rate and the step kernel's average duration.

Also reported (extra fields): the same measurement for the FP32-storage form,
full solves to 1e-4 relative KKT for both forms (time from after the upload
to the terminating evaluation, objectives against each other), and the CPU
baseline: the oracle's literal restatement of the reference on the same
instance, a bounded number of iterations on the host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-solve] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "pdot.jl_amd"))

METRIC = "time-to-1e-4 rel KKT at 10k×10k, FP32/learned vs FP64; iters/s; HBM GB/s %"
PEAK_HBM_GBS = 8000.0


def algorithmic_bytes_per_step(n: int, m: int, elem: int, matrix_streams: int = 4) -> int:
    """Bytes one PDHG iteration moves on average, plus the row/column vectors
    (u, v, r, c and their averages read and written, a and b read: about 10
    doubles per row and per column).  A pass reads C and X and writes X+ (3
    n*m element streams); the running average is read and written (2 more)
    by every second pass only, which also applies the update the pass before
    it deferred: 4 streams per iteration on average (3 without a stored C,
    the point-cloud form).  With deferred averaging off, 5."""
    return matrix_streams * n * m * elem + 10 * 8 * (n + m)


def measured_traffic(n: int, m: int, form: str):
    """HBM bytes per step-kernel launch measured with rocprofv3 FETCH_SIZE /
    WRITE_SIZE (scaled by a copy kernel of known size) for this workload, as
    recorded in profiles/hbm_traffic_step_kernel_gaussian_shift_10k.md; None
    for other sizes."""
    path = os.path.join(REPO, "profiles", "hbm_traffic_step_kernel_gaussian_shift_10k.md")
    if (n, m) != (10000, 10000) or not os.path.exists(path):
        return None
    text = open(path).read()
    block = text.split("```json", 1)[1].split("```", 1)[0]
    return json.loads(block)[form]["traffic_bytes"]


def measured_batch_traffic():
    """HBM bytes per batched step-kernel launch (every instance active) of the
    512-instance workload, from profiles/hbm_traffic_bstep_kernel_batched_uniform_512.md."""
    path = os.path.join(REPO, "profiles", "hbm_traffic_bstep_kernel_batched_uniform_512.md")
    if not os.path.exists(path):
        return None
    text = open(path).read()
    return json.loads(text.split("```json", 1)[1].split("```", 1)[0])["traffic_bytes"]


def measure_steps(dev, steps: int, warmup: int):
    from pdot_amd import _lib
    dev.set_start(policy=_lib.POLICY_ADAPTIVE, step_size=1.0, primal_weight=1.0)
    if warmup:
        dev.time_steps(warmup)
    ks, tt = dev.time_steps(steps)
    st = dev.get_state()
    return ks, tt, st


def cpu_baseline(prob, iters: int):
    """The reference algorithm restated on host arrays (oracle literal mode:
    stores G and its average like the reference, 8 n*m FP64 arrays),
    OpenMP across the host cores, timed inside take_step! only."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    import pdot_amd as P
    params = P.default_parameters(1e-4, iteration_limit=iters + 1)
    t0 = time.time()
    r = O.optimize(params, prob, mode=O.LITERAL)
    wall = time.time() - t0
    st = r["stats"][-1]
    steps = int(r["take_step_calls"])
    basic = st["time_spent_doing_basic_algorithm"]
    return {"value": steps / basic, "unit": "iters/s", "cores": O.num_threads(), "kind": "port",
            "sample": f"{steps} take_step! iterations of the literal FP64 restatement on the same "
                      f"{prob.num_row}x{prob.num_col} instance (evaluations excluded; wall {wall:.1f}s incl. "
                      f"{len(r['stats'])} evaluations)"}


def solve_to_tol(prob, storage, tol, dev=None, do_round=True, primal_importance=None):
    from dataclasses import replace
    import pdot_amd as P
    prof = P.SolveProfile()
    params = P.default_parameters(tol, verbosity=0, record_iteration_stats=False)
    if primal_importance is not None:
        params = replace(params, primal_importance=float(primal_importance))
    out = P.optimize(params, prob, storage=storage, solver=dev, profile=prof)
    ci = out.iteration_stats[-1].convergence_information[0]
    res = {"seconds": prof.solve_sec, "output_copy_sec": prof.output_sec, "iterations": out.iteration_count,
           "termination": out.termination_string, "objective": ci.primal_objective,
           "rel_kkt": P.relative_kkt_error(ci), "evaluations": prof.evaluations, "restarts": prof.restarts}
    if dev is not None and do_round:
        # output post-processing: feasibility rounding of the resident solution on the device
        # (four reads of X, one FP64 write) plus the device-to-host copy of the rounded plan
        t0 = time.time()
        Xr = dev.round(dev.output_point)
        res["rounding_sec_incl_copy"] = time.time() - t0
        if hasattr(prob, "cost_matrix"):
            res["rounded_objective"] = float((prob.cost_matrix * Xr).sum())
        del Xr
    return res


def bench_l1_mixture(args) -> int:
    """30000 x 30000 L1-cost mixture, FP32 storage (C, X, X_avg 3.6 GB each)
    with FP64 duals; the FP64 form on the same instance gives the objective
    error of the FP32-storage result."""
    from pdot_amd import _lib, instances
    n = m = 30000 if args.n == 10000 else args.n
    t0 = time.time()
    prob = instances.l1_mixture_3d(n)
    gen = time.time() - t0
    out = {"metric": "30k x 30k L1 mixture: FP32-storage vs FP64 solve to 1e-4", "unit": "s",
           "higher_is_better": False, "data": "synthetic", "n_gpus": 1,
           "config": {"workload": f"l1_mixture_3d {n}x{m}", "n": n, "m": m}, "generate_sec": gen}
    for storage, name, elem in ((_lib.STORAGE_F32, "f32", 4), (_lib.STORAGE_F64, "f64", 8)):
        dev = _lib.DeviceSolver(n, m, storage=storage)
        t0 = time.time()
        dev.upload(prob.cost_matrix, prob.source_distribution, prob.target_distribution)
        up = time.time() - t0
        ks, tt, st = measure_steps(dev, args.steps, args.warmup)
        alg = algorithmic_bytes_per_step(n, m, elem)
        k_avg = ks / args.steps * 1e-3
        res = {"upload_sec": up, "ms_per_step": tt / args.steps, "step_kernel_GBs": alg / k_avg / 1e9,
               "device_GB": dev.device_bytes() / 1e9}
        if not args.no_solve:
            res["solve"] = solve_to_tol(prob, storage, 1e-4, dev=dev, do_round=False)
        out[name] = res
        dev.close()
    # point-cloud cost form, FP32 storage: C never stored, 4 n*m streams per step
    pts = instances.l1_mixture_3d_points(n)
    dev = _lib.DeviceSolver(n, m, storage=_lib.STORAGE_F32)
    t0 = time.time()
    dev.upload_points(pts)
    up = time.time() - t0
    ks, tt, st = measure_steps(dev, args.steps, args.warmup)
    alg = algorithmic_bytes_per_step(n, m, 4, matrix_streams=3)
    res = {"upload_sec": up, "ms_per_step": tt / args.steps, "step_kernel_GBs": alg / (ks / args.steps * 1e-3) / 1e9,
           "device_GB": dev.device_bytes() / 1e9}
    if not args.no_solve:
        res["solve"] = solve_to_tol(pts, _lib.STORAGE_F32, 1e-4, dev=dev, do_round=False)
    out["f32_point_cloud"] = res
    dev.close()
    out["roofline"] = {"bound": "hbm", "achieved": out["f32"]["step_kernel_GBs"], "peak": PEAK_HBM_GBS, "unit": "GB/s",
                       "frac": out["f32"]["step_kernel_GBs"] / PEAK_HBM_GBS, "traffic": None,
                       "kernel": "step_kernel<float> (FP32 storage)",
                       "bytes_per_launch": algorithmic_bytes_per_step(n, m, 4)}
    if not args.no_solve:
        out["value"] = out["f32"]["solve"]["seconds"]
        out["objective_rel_err_f32_vs_f64"] = (abs(out["f32"]["solve"]["objective"] - out["f64"]["solve"]["objective"])
                                               / abs(out["f64"]["solve"]["objective"]))
    print(json.dumps(out))
    return 0


def bench_batched(args) -> int:
    """512 instances, n and m uniform in [64, 2048], FP32 storage, every
    instance solved to 1e-4 relative KKT in one batch."""
    import numpy as np
    import pdot_amd as P
    from pdot_amd import _lib, instances
    from pdot_amd.batch import BatchProfile, autotune, optimize_batch
    t0 = time.time()
    probs = instances.batched_uniform(512)
    gen = time.time() - t0
    ns, ms = [p.num_row for p in probs], [p.num_col for p in probs]
    b = _lib.BatchSolver(ns, ms, storage=_lib.STORAGE_F32)
    t0 = time.time()
    for k, p in enumerate(probs):
        b.upload(k, p.cost_matrix, p.source_distribution, p.target_distribution)
    up = time.time() - t0
    tune = autotune(b, steps=4)
    b.set_start(_lib.POLICY_ADAPTIVE, 2.0, 1.0, 1.0, 1.0)
    b.time_steps(1)  # the first pass after a start has an empty average (untimed)
    # an even count: the passes that defer the averages' update alternate with those applying it
    steps = max(2, args.steps // 8 * 2)
    ks, tt = b.time_steps(steps)
    total_elems = sum(n * m for n, m in zip(ns, ms))
    alg = 4 * 4 * total_elems + 80 * (sum(ns) + sum(ms))  # 3 and 5 element streams alternately
    res = {"metric": "512-instance batch (n, m in [64, 2048]) solved to 1e-4, FP32 storage", "unit": "s",
           "higher_is_better": False, "data": "synthetic", "n_gpus": 1,
           "config": {"workload": "batched_uniform 512 x [64..2048]^2", "instances": 512,
                      "elements": total_elems, "storage": "f32"},
           "generate_sec": gen, "upload_sec": up, "autotune": tune["best"],
           "pass_ms_all_active": tt / steps, "step_kernel_GBs": alg / (ks / steps * 1e-3) / 1e9,
           "bytes_per_pass": alg}
    res["roofline"] = {"bound": "hbm", "achieved": res["step_kernel_GBs"], "peak": PEAK_HBM_GBS, "unit": "GB/s",
                       "frac": res["step_kernel_GBs"] / PEAK_HBM_GBS, "traffic": measured_batch_traffic(),
                       "kernel": "bstep_kernel<float> (every instance active)", "bytes_per_launch": alg,
                       "avg_launch_us": ks / steps * 1e3}
    if not args.no_solve:
        prof = BatchProfile()
        outs = optimize_batch(P.default_parameters(1e-4, record_iteration_stats=False), probs, batch=b,
                              use_graphs=True, profile=prof)
        iters = np.array([o.iteration_count for o in outs])
        optimal = sum(o.termination_reason == P.TerminationReason.TERMINATION_REASON_OPTIMAL for o in outs)
        res.update({"value": prof.solve_sec, "solve_sec": prof.solve_sec, "optimal": int(optimal),
                    "instances_per_s": 512 / prof.solve_sec, "iterations_median": float(np.median(iters)),
                    "iterations_max": int(iters.max()), "passes_launched": prof.passes_launched,
                    "evaluations": prof.evaluations})
        # the same batch with the initial primal weight calibrated for the learned warm start
        # (primal_importance; the reference's default, used above, is 1.0)
        from dataclasses import replace
        from pdot_amd import learned
        w0 = learned.load_initial_primal_weight()
        if w0 is not None:
            prof2 = BatchProfile()
            outs2 = optimize_batch(replace(P.default_parameters(1e-4, record_iteration_stats=False),
                                           primal_importance=w0), probs, batch=b, use_graphs=True, profile=prof2)
            it2 = np.array([o.iteration_count for o in outs2])
            obj = np.array([o.iteration_stats[-1].convergence_information[0].primal_objective for o in outs])
            obj2 = np.array([o.iteration_stats[-1].convergence_information[0].primal_objective for o in outs2])
            res["with_primal_importance"] = {
                "primal_importance": w0, "solve_sec": prof2.solve_sec,
                "optimal": int(sum(o.termination_reason == P.TerminationReason.TERMINATION_REASON_OPTIMAL
                                   for o in outs2)),
                "iterations_median": float(np.median(it2)), "iterations_max": int(it2.max()),
                "objective_rel_diff_max": float(np.max(np.abs(obj2 - obj) / np.abs(obj)))}
    b.close()
    print(json.dumps(res))
    return 0


def bench_precision(args) -> int:
    """FP32 storage against the FP64 oracle at 10000 x 10000 (gaussian shift):
    the FP64 solve tightened to 1e-8 and the FP32-storage solve to 1.5e-8,
    each with the progress table printed every 50 evaluations.  About 15
    minutes on one MI355X."""
    import numpy as np
    import pdot_amd as P
    from pdot_amd import _lib, instances
    n = args.n
    prob = instances.gaussian_shift(n)
    out = {"metric": "objective of FP32 storage solved to 1.5e-8 vs the FP64 solve to 1e-8 (relative)",
           "unit": "relative", "higher_is_better": False, "data": "synthetic", "n_gpus": 1,
           "config": {"workload": f"gaussian_shift {n}x{n}", "n": n, "m": n}}
    for name, storage, tol in (("f64_oracle", _lib.STORAGE_F64, 1e-8), ("f32", _lib.STORAGE_F32, 1.5e-8)):
        params = P.default_parameters(tol, verbosity=3, record_iteration_stats=True)
        t0 = time.time()
        sol = P.optimize(params, prob, storage=storage)
        wall = time.time() - t0
        ci = sol.iteration_stats[-1].convergence_information[0]
        out[name] = {"tolerance": tol, "termination": sol.termination_string, "iterations": sol.iteration_count,
                     "seconds": wall, "rel_kkt": P.relative_kkt_error(ci),
                     "objective": float(np.sum(prob.cost_matrix * sol.primal_solution))}
        sys.stdout.flush()
    out["value"] = abs(out["f32"]["objective"] - out["f64_oracle"]["objective"]) / abs(out["f64_oracle"]["objective"])
    print(json.dumps(out))
    return 0


def bench_warm_start(args) -> int:
    """Held-out 4096 x 4096 Gaussian-mixture instances (FP32 storage): time to
    1e-4 relative KKT from the zero start and through the learned warm start
    (committed network weights, KKT acceptance test and zero-start fallback), the
    prediction time included."""
    import numpy as np
    import pdot_amd as P
    from pdot_amd import _lib, instances, learned
    net = learned.WarmStartNet(learned.load_params())
    rows = []
    for k in range(args.instances):
        inst = instances.gmm_dirichlet(4096, seed=900200 + k)
        p = inst.problem
        prm = P.default_parameters(1e-4, record_iteration_stats=False)
        t0 = time.time()
        zero = P.optimize(prm, p, storage=_lib.STORAGE_F32)
        tz = time.time() - t0
        t0 = time.time()
        warm, info = learned.solve_warm(prm, p, inst.x, inst.y, net, storage=_lib.STORAGE_F32)
        tw = time.time() - t0
        rows.append({"seed": 900200 + k, "zero_iterations": zero.iteration_count, "zero_sec": tz,
                     "warm_iterations": warm.iteration_count, "warm_sec": tw, "duals_used": info.used_duals,
                     "initial_primal_weight": info.primal_weight, "fell_back": info.fell_back,
                     "predict_sec": info.predict_sec, "warm_termination": warm.termination_string,
                     "objective_rel_diff": abs(warm.iteration_stats[-1].convergence_information[0].primal_objective
                                               - zero.iteration_stats[-1].convergence_information[0].primal_objective)
                                           / abs(zero.iteration_stats[-1].convergence_information[0].primal_objective)})
        print(f"# instance {k}: zero {zero.iteration_count} it {tz:.2f}s, warm {warm.iteration_count} it {tw:.2f}s",
              file=sys.stderr, flush=True)
    net.close()
    tz = sum(r["zero_sec"] for r in rows)
    tw = sum(r["warm_sec"] for r in rows)
    print(json.dumps({"metric": "learned warm start: time-to-1e-4 on held-out 4096x4096 GMM instances (FP32 storage)",
                      "value": tz / tw, "unit": "speedup (zero-start time / warm-start time, incl. prediction)",
                      "higher_is_better": True, "data": "synthetic", "n_gpus": 1,
                      "config": {"workload": "gmm_dirichlet 4096x4096", "instances": len(rows), "storage": "f32"},
                      "zero_sec_total": tz, "warm_sec_total": tw,
                      "iteration_ratio_median": float(np.median([r["zero_iterations"] / max(1, r["warm_iterations"])
                                                                 for r in rows])),
                      "rows": rows}))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=10000)
    ap.add_argument("--no-solve", action="store_true", help="skip the time-to-tolerance solves")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-iters", type=int, default=80, help="iteration limit of the CPU baseline sample (about 15 s)")
    ap.add_argument("--instances", type=int, default=8, help="held-out instances of the gmm_warm_start workload")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run a plain device copy of this many MiB (counter calibration)")
    ap.add_argument("--workload", default="gaussian_shift",
                    choices=("gaussian_shift", "l1_mixture_3d", "batched_uniform", "gmm_warm_start", "precision"),
                    help="gaussian_shift: the headline 10k x 10k measurement; the others measure the 30k x 30k "
                         "FP32-storage stress case, the 512-instance batch, the learned warm start and (about "
                         "15 minutes) FP32 storage against the FP64 1e-8 solve at 10k x 10k")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the library has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        return 2
    if args.workload == "l1_mixture_3d":
        return bench_l1_mixture(args)
    if args.workload == "batched_uniform":
        return bench_batched(args)
    if args.workload == "precision":
        return bench_precision(args)
    if args.workload == "gmm_warm_start":
        return bench_warm_start(args)

    from pdot_amd import _lib, instances

    n = m = args.n
    prob = instances.gaussian_shift(n)
    results = {}
    devs = {}
    for storage, name, elem in ((_lib.STORAGE_F64, "f64", 8), (_lib.STORAGE_F32, "f32", 4)):
        dev = _lib.DeviceSolver(n, m, storage=storage)
        t0 = time.time()
        dev.upload(prob.cost_matrix, prob.source_distribution, prob.target_distribution)
        up = time.time() - t0
        ks, tt, st = measure_steps(dev, args.steps, args.warmup)
        alg = algorithmic_bytes_per_step(n, m, elem)
        k_avg = ks / args.steps * 1e-3
        results[name] = {
            "ms_per_step": tt / args.steps, "iters_per_s": args.steps / (tt * 1e-3),
            "step_kernel_us": k_avg * 1e6, "step_kernel_GBs": alg / k_avg / 1e9,
            "iteration_GBs": alg / (tt / args.steps * 1e-3) / 1e9, "bytes_per_step": alg,
            "launch": dev.get_launch(), "numerical_error": int(st.numerical_error),
            "upload_sec": up, "upload_GBs_pcie": 8 * n * m / up / 1e9,
        }
        devs[name] = dev

    # point-cloud cost form (C formed in registers, never stored): 4 n*m streams per step
    points = {}
    pts = instances.gaussian_shift_points(n, m)
    for storage, name, elem in ((_lib.STORAGE_F64, "f64", 8), (_lib.STORAGE_F32, "f32", 4)):
        dev = _lib.DeviceSolver(n, m, storage=storage)
        dev.upload_points(pts)
        ks, tt, st = measure_steps(dev, args.steps, args.warmup)
        alg = algorithmic_bytes_per_step(n, m, elem, matrix_streams=3)
        k_avg = ks / args.steps * 1e-3
        points[name] = {"ms_per_step": tt / args.steps, "iters_per_s": args.steps / (tt * 1e-3),
                        "step_kernel_us": k_avg * 1e6, "step_kernel_GBs": alg / k_avg / 1e9, "bytes_per_step": alg,
                        "device_GB": dev.device_bytes() / 1e9}
        if not args.no_solve:
            points[name]["time_to_1e-4"] = solve_to_tol(pts, storage, 1e-4, dev=dev)
        dev.close()

    solves = {}
    if not args.no_solve:
        for name, storage in (("f64", _lib.STORAGE_F64), ("f32", _lib.STORAGE_F32)):
            solves[name] = solve_to_tol(prob, storage, 1e-4, dev=devs[name])
        solves["speedup_f32_vs_f64"] = solves["f64"]["seconds"] / solves["f32"]["seconds"]
        solves["objective_rel_diff_f32_vs_f64"] = (abs(solves["f32"]["objective"] - solves["f64"]["objective"])
                                                   / abs(solves["f64"]["objective"]))
        # the learned starting state's initial primal weight (calibrated on Gaussian-mixture
        # solves; the reference's default primal_importance is 1.0)
        from pdot_amd import learned
        w0 = learned.load_initial_primal_weight()
        if w0 is not None:
            for name, storage in (("f64", _lib.STORAGE_F64), ("f32", _lib.STORAGE_F32)):
                r = solve_to_tol(prob, storage, 1e-4, dev=devs[name], do_round=False, primal_importance=w0)
                r["primal_importance"] = w0
                solves[f"{name}_calibrated_primal_weight"] = r
            solves["speedup_calibrated_f64_vs_f64"] = solves["f64"]["seconds"] / solves["f64_calibrated_primal_weight"]["seconds"]
            solves["speedup_calibrated_f32_vs_f64"] = solves["f64"]["seconds"] / solves["f32_calibrated_primal_weight"]["seconds"]
    copy_probe = None
    if args.copy_probe_mib:
        nbytes = args.copy_probe_mib << 20
        ms = devs["f64"].copy_probe(nbytes, reps=5)
        copy_probe = {"bytes_per_copy": 2 * nbytes, "ms": ms, "GBs": 2 * nbytes / (ms * 1e-3) / 1e9,
                      "launches": 6}
    for d in devs.values():
        d.close()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(prob, args.cpu_iters)

    f64 = results["f64"]
    line = {
        "metric": METRIC,
        "value": f64["iters_per_s"],
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": f64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"gaussian_shift {n}x{m} FP64 (squared Euclidean / max, U(0,1) masses, seed 2000)",
                   "n": n, "m": m, "storage": "f64", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": f64["step_kernel_GBs"], "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": f64["step_kernel_GBs"] / PEAK_HBM_GBS, "traffic": measured_traffic(n, m, "f64"),
                     "kernel": "step_kernel<double>", "bytes_per_launch": f64["bytes_per_step"],
                     "avg_launch_us": f64["step_kernel_us"]},
        "cpu_baseline": cpu,
        "iteration_GBs": f64["iteration_GBs"],
        "f32_storage": results["f32"],
        "time_to_1e-4": solves or None,
        "point_cloud_cost": points,
    }
    if copy_probe:
        line["copy_probe"] = copy_probe
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
