"""Train the warm-start network and measure what it gains.

    python -m pdot_amd.train_warm_start --out weights.npz --report report.json

1. Training data: 2-D Gaussian-mixture instances (3-8 components, means
   U[0,1]^2, sigma U[0.02, 0.15], Dirichlet(1) weights and masses, squared
   Euclidean cost) of random sizes, solved by the batched plain solver
   (zero start) to a tight tolerance; their duals, normalised for the
   (u + t, v - t) invariance, are the targets.
2. Training: Adam on the GPU through pdot_mlp, one instance per step.
3. Calibration of the initial primal weight (calibrate_primal_weight): the
   weight of a small grid with the fewest total iterations over zero-start
   solves of separate calibration instances at n = m = 4096; stored beside
   the network weights (`--calibrate-only` redoes just this step).
4. Held-out evaluation on instances of the same family at n = m = 4096 with
   seeds disjoint from training: iterations and time to 1e-4 from the zero
   start and from the predicted duals (same solver, same parameters; only
   the starting iterate differs).
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import _lib, default_parameters, instances, learned
from .batch import optimize_batch
from .solve_log import TerminationReason
from .solver import SolveProfile, optimize, relative_kkt_error

TRAIN_SEED0 = 100000
CALIB_SEED0 = 200000
EVAL_SEED0 = 900000
PRIMAL_WEIGHT_GRID = (0.05, 0.1, 0.2, 0.3, 0.5, 1.0)


def calibrate_primal_weight(count=16, n=4096, grid=PRIMAL_WEIGHT_GRID, tol=1e-4, storage=_lib.STORAGE_F32, log=print):
    """The initial primal weight (the reference's primal_importance) that
    minimises the total iteration count to `tol` over `count` calibration
    instances of the family (seeds CALIB_SEED0 + k, disjoint from training
    and evaluation), each solved from the zero start with every weight of
    `grid`.  Returns (best weight, {weight: [iterations per instance]})."""
    from dataclasses import replace
    table = {float(w): [] for w in grid}
    for k in range(count):
        p = instances.gmm_dirichlet(n, seed=CALIB_SEED0 + k).problem
        dev = _lib.DeviceSolver(n, n, storage=storage)
        dev.upload(p.cost_matrix, p.source_distribution, p.target_distribution)
        for w in grid:
            params = replace(default_parameters(tol, record_iteration_stats=False), primal_importance=float(w))
            out = optimize(params, p, storage=storage, solver=dev)
            ok = out.termination_reason == TerminationReason.TERMINATION_REASON_OPTIMAL
            table[float(w)].append(out.iteration_count if ok else 10 ** 9)
        dev.close()
        log(f"calibration {k}: " + " ".join(f"{w:g}:{table[float(w)][-1]}" for w in grid))
    best = min(table, key=lambda w: (sum(table[w]), w))
    return best, table


def generate_training_data(count, n_min, n_max, tol, seed, batch_size=512, iteration_limit=20000, log=print):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(n_min, n_max + 1, size=(count, 2))
    samples, kept, t0 = [], 0, time.time()
    for start in range(0, count, batch_size):
        idx = range(start, min(count, start + batch_size))
        insts = [instances.gmm_dirichlet(int(sizes[k, 0]), int(sizes[k, 1]), seed=TRAIN_SEED0 + seed * 100000 + k)
                 for k in idx]
        outs = optimize_batch(default_parameters(tol, iteration_limit=iteration_limit, record_iteration_stats=False),
                              [i.problem for i in insts], storage=_lib.STORAGE_F64)
        for inst, out in zip(insts, outs):
            if out.termination_reason != TerminationReason.TERMINATION_REASON_OPTIMAL:
                continue
            p = inst.problem
            samples.append(learned.make_sample(inst.x, inst.y, p.source_distribution, p.target_distribution,
                                               out.dual_source_solution, out.dual_target_solution))
            kept += 1
        log(f"data: {start + len(idx)}/{count} solved, {kept} kept, {time.time() - t0:.1f}s")
    return samples


def evaluate(net, count, n, tols=(1e-4,), storage=_lib.STORAGE_F64, log=print):
    """Held-out comparison: iterations and time to each tolerance from the
    zero start and through learned.solve_warm (prediction, KKT acceptance test,
    fallback), same solver and parameters.  Warm solves are capped at 20x the
    zero start's iterations (then reported as not reaching the tolerance)."""
    from dataclasses import replace
    rows = []
    for k in range(count):
        inst = instances.gmm_dirichlet(n, seed=EVAL_SEED0 + k)
        p = inst.problem
        row = {"seed": EVAL_SEED0 + k}
        dev = _lib.DeviceSolver(n, n, storage=storage)
        dev.upload(p.cost_matrix, p.source_distribution, p.target_distribution)
        t0 = time.time()
        u0, v0 = net.predict_duals(inst.x, inst.y, p.source_distribution, p.target_distribution)
        row["predict_sec"] = time.time() - t0
        for tol in tols:
            params = default_parameters(tol, record_iteration_stats=False)
            prof = SolveProfile()
            out = optimize(params, p, storage=storage, solver=dev, profile=prof)
            ci = out.iteration_stats[-1].convergence_information[0] if out.iteration_stats else None
            row[f"zero_{tol:g}"] = {"iterations": out.iteration_count, "seconds": prof.solve_sec,
                                     "termination": out.termination_string}
            log(f"eval {k} tol {tol:g} zero: {out.iteration_count} it {prof.solve_sec:.2f}s {out.termination_string}")
            cap = replace(params, termination_criteria=replace(params.termination_criteria,
                                                               iteration_limit=max(1000, 20 * out.iteration_count)))
            prof = SolveProfile()
            out = optimize(cap, p, storage=storage, warm_start=(u0, v0), solver=dev, profile=prof)
            row[f"warm_{tol:g}"] = {"iterations": out.iteration_count, "seconds": prof.solve_sec,
                                     "termination": out.termination_string}
            log(f"eval {k} tol {tol:g} warm: {out.iteration_count} it {prof.solve_sec:.2f}s {out.termination_string}")
        dev.close()
        rows.append(row)
    summary = {}
    for tol in tols:
        z = np.array([r[f"zero_{tol:g}"]["iterations"] for r in rows], dtype=float)
        w = np.array([r[f"warm_{tol:g}"]["iterations"] for r in rows], dtype=float)
        zt = np.array([r[f"zero_{tol:g}"]["seconds"] for r in rows])
        wt = np.array([r[f"warm_{tol:g}"]["seconds"] + r["predict_sec"] for r in rows])
        summary[f"{tol:g}"] = {"iterations_zero_median": float(np.median(z)), "iterations_warm_median": float(np.median(w)),
                               "iteration_ratio_median": float(np.median(z / np.maximum(w, 1))),
                               "time_zero_total": float(zt.sum()), "time_warm_total_incl_predict": float(wt.sum()),
                               "time_ratio_total": float(zt.sum() / wt.sum()),
                               "warm_optimal": int(sum(r[f"warm_{tol:g}"]["termination"] == "OPTIMAL" for r in rows))}
    return rows, summary


def calibration_meta(best, table) -> dict:
    weights = sorted(table)
    return {"initial_primal_weight": float(best),
            "calibration_grid": np.array(weights, dtype=np.float64),
            "calibration_iterations": np.array([table[w] for w in weights], dtype=np.int64),
            "calibration_seed0": CALIB_SEED0}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=1024, help="training instances")
    ap.add_argument("--n-min", type=int, default=256)
    ap.add_argument("--n-max", type=int, default=768)
    ap.add_argument("--data-tol", type=float, default=1e-5)
    ap.add_argument("--epochs", type=int, default=25)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eval", type=int, default=8, help="held-out instances")
    ap.add_argument("--eval-n", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=learned.DEFAULT_WEIGHTS)
    ap.add_argument("--report", default=None)
    ap.add_argument("--calibrate", type=int, default=16, help="calibration instances for the initial primal weight")
    ap.add_argument("--calibrate-only", action="store_true",
                    help="keep the network weights in --out and only (re)calibrate the initial primal weight")
    args = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)  # noqa: E731

    if args.calibrate_only:
        params, meta = learned.load_params(args.out), learned.load_meta(args.out)
        best, table = calibrate_primal_weight(args.calibrate, log=log)
        meta.update(calibration_meta(best, table))
        learned.save_params(args.out, params, meta)
        log(json.dumps({"initial_primal_weight": best, "total_iterations": {w: sum(v) for w, v in table.items()}}))
        return 0

    t0 = time.time()
    samples = generate_training_data(args.train, args.n_min, args.n_max, args.data_tol, args.seed, log=log)
    data_sec = time.time() - t0
    rng = np.random.default_rng(args.seed)
    order = rng.permutation(len(samples))
    nval = max(1, len(samples) // 10)
    val = [samples[i] for i in order[:nval]]
    trn = [samples[i] for i in order[nval:]]
    net = learned.WarmStartNet(seed=args.seed)
    zero = float(np.mean([np.mean(s.target.astype(np.float64) ** 2) for s in val]))
    t0 = time.time()
    hist = learned.train(net, trn, epochs=args.epochs, lr=args.lr, seed=args.seed, log=log)
    train_sec = time.time() - t0
    val_loss = learned.evaluate_loss(net, val)
    log(f"validation mse {val_loss:.4e} (predicting zero: {zero:.4e})")
    params = net.get_params()
    meta = {"train_instances": len(trn), "val_mse": val_loss, "epochs": args.epochs}
    if args.calibrate > 0:
        meta.update(calibration_meta(*calibrate_primal_weight(args.calibrate, log=log)))
    learned.save_params(args.out, params, meta)
    rows, summary = evaluate(net, args.eval, args.eval_n, log=log) if args.eval > 0 else ([], {})
    report = {"data_sec": data_sec, "train_sec": train_sec, "train_instances": len(trn), "val_instances": len(val),
              "train_loss_history": hist, "val_mse": val_loss, "val_mse_zero_prediction": zero,
              "eval_n": args.eval_n, "eval": rows, "summary": summary}
    log(json.dumps(summary))
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
