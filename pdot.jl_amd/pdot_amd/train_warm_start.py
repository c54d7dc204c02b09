"""Train the warm-start network and measure what it gains.

    python -m pdot_amd.train_warm_start --out weights.npz --report report.json

1. Weight data: Gaussian-mixture instances of the c5 family (2-D mixtures
   with 3-8 components, means U[0,1]^2, sigma U[0.02, 0.15], Dirichlet(1)
   component weights and masses, squared Euclidean cost) at n, m log-uniform
   in [--n-min, --n-max] plus --grid-4096 instances at 4096 x 4096, each
   solved to 1e-4 from the zero start by the batched solver (FP32 storage) at
   every initial primal weight of WEIGHT_GRID (runs capped at 3x the
   instance's best so far).  An instance's target is the log of its best
   weight, refined by a parabola through the log iteration counts around the
   grid minimum.
2. Calibrated constant: the grid weight with the fewest total iterations over
   the training instances (the learned start's fallback weight and the
   baseline the head must beat).
3. Weight head: Adam on the GPU through pdot_mlp (pdot_mlp_weight_loss_grad),
   head only (the dual network and its first layer stay fixed), with the
   validation split's loss selecting the epoch.
4. Held-out evaluation at n = m = 4096 (seeds disjoint from training): time
   and iterations to 1e-4 from the zero start with the reference's weight 1.0,
   the calibrated constant and the head's per-instance weight.  Every solve
   prints a progress line.

The dual network (predicting (u, v)) is trained with --train N (duals of
plain solves at n, m in [--dual-n-min, --dual-n-max]); 0 keeps the dual
parameters already in --out.
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
WEIGHT_SEED0 = 300000
HELD_OUT_SEED0 = 900200  # the held-out instances the benchmark reports
# initial primal weights of the per-instance grid, in the order they are run
# (the middle first, so that the caps of the later runs are tight)
WEIGHT_GRID = (0.2, 0.13, 0.32, 0.08, 0.5, 0.05, 0.8, 0.03)


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


def weight_grid_data(count, n_min, n_max, seed0, grid=WEIGHT_GRID, tol=1e-4, batch_size=256, cap_factor=3.0,
                     fixed_n=None, first_cap=32768, log=print):
    """Iterations to `tol` from the zero start at every initial primal weight
    of `grid` for `count` GMM instances (seeds seed0 + k; sizes log-uniform in
    [n_min, n_max], or fixed_n x fixed_n), FP32 storage, batched.  A run at a
    later weight stops at cap_factor times the instance's best so far (rounded
    up to the evaluation frequency; the first weight's runs stop at first_cap)
    and counts as not reaching the tolerance.
    Returns {"seed", "n", "m", "grid", "iterations" (count x len(grid), -1 where
    capped or not OPTIMAL)}."""
    from dataclasses import replace
    rng = np.random.default_rng(seed0)
    if fixed_n:
        sizes = np.full((count, 2), int(fixed_n))
    else:
        sizes = np.exp(rng.uniform(np.log(n_min), np.log(n_max + 1), size=(count, 2))).astype(np.int64)
    table = np.full((count, len(grid)), -1, dtype=np.int64)
    t0 = time.time()
    for start in range(0, count, batch_size):
        idx = list(range(start, min(count, start + batch_size)))
        insts = [instances.gmm_dirichlet(int(sizes[k, 0]), int(sizes[k, 1]), seed=seed0 + k) for k in idx]
        probs = [i.problem for i in insts]
        b = _lib.BatchSolver([p.num_row for p in probs], [p.num_col for p in probs], storage=_lib.STORAGE_F32)
        for j, p in enumerate(probs):
            b.upload(j, p.cost_matrix, p.source_distribution, p.target_distribution)
        best = np.full(len(idx), np.inf)
        for gi, w in enumerate(grid):
            base = replace(default_parameters(tol, record_iteration_stats=False), primal_importance=float(w))
            plist = []
            for j in range(len(idx)):
                prm = replace(base, termination_criteria=replace(base.termination_criteria))
                cap = first_cap if not np.isfinite(best[j]) else int(np.ceil(cap_factor * best[j] / 128) * 128)
                prm.termination_criteria.iteration_limit = cap
                plist.append(prm)
            t1 = time.time()
            outs = optimize_batch(plist, probs, batch=b, use_graphs=True)
            for j, out in enumerate(outs):
                if out.termination_reason == TerminationReason.TERMINATION_REASON_OPTIMAL:
                    table[idx[j], gi] = out.iteration_count
                    best[j] = min(best[j], out.iteration_count)
            log(f"weight data: instances {idx[0]}-{idx[-1]} weight {w:g}: "
                f"{int((table[idx, gi] >= 0).sum())}/{len(idx)} optimal, median "
                f"{np.median(table[idx, gi][table[idx, gi] >= 0]) if (table[idx, gi] >= 0).any() else -1:.0f} it, "
                f"{time.time() - t1:.1f}s (total {time.time() - t0:.0f}s)")
        b.close()
    return {"seed": seed0 + np.arange(count), "n": sizes[:, 0], "m": sizes[:, 1],
            "grid": np.array(grid, dtype=np.float64), "iterations": table}


def best_log_weight(grid, iters) -> float:
    """The log of an instance's best initial weight: the grid minimum of the
    iteration count, refined by a parabola through log(iterations) at it and
    its two neighbours in log(weight) (kept within the neighbours)."""
    order = np.argsort(grid)
    lw = np.log(np.asarray(grid, dtype=np.float64)[order])
    it = np.asarray(iters, dtype=np.float64)[order]
    it = np.where(it > 0, it, np.inf)
    i = int(np.argmin(it))
    if not np.isfinite(it[i]):
        return float("nan")
    if 0 < i < len(lw) - 1 and np.isfinite(it[i - 1]) and np.isfinite(it[i + 1]):
        x, y = lw[i - 1:i + 2], np.log(it[i - 1:i + 2])
        a, b_, _ = np.polyfit(x, y, 2)
        if a > 0:
            return float(np.clip(-b_ / (2 * a), x[0], x[2]))
    return float(lw[i])


def calibrated_constant(data) -> float:
    """The grid weight with the fewest total iterations over the data
    (capped or failed runs count as 3x the instance's best)."""
    it = data["iterations"].astype(np.float64)
    best = np.where(it > 0, it, np.inf).min(axis=1)
    it = np.where(it > 0, it, 3 * best[:, None])
    ok = np.isfinite(best)
    tot = it[ok].sum(axis=0)
    return float(data["grid"][int(np.argmin(tot))])


def weight_samples(data):
    """(features, ns, nt, log-weight target, instance index) per instance with
    a finite target."""
    out = []
    for k in range(len(data["seed"])):
        t = best_log_weight(data["grid"], data["iterations"][k])
        if not np.isfinite(t):
            continue
        x, y, a, b = instances.gmm_dirichlet_points(int(data["n"][k]), int(data["m"][k]), seed=int(data["seed"][k]))
        out.append((learned.features(x, y, a, b), len(x), len(y), t, k))
    return out


def grid_iterations(data, k, weight) -> float:
    """Iterations of instance k at the grid weight nearest `weight` (in log
    scale); a capped run counts as 3x the instance's best."""
    g = np.asarray(data["grid"], dtype=np.float64)
    it = data["iterations"][k].astype(np.float64)
    best = np.where(it > 0, it, np.inf).min()
    j = int(np.argmin(np.abs(np.log(g) - np.log(weight))))
    return float(it[j]) if it[j] > 0 else 3.0 * best


SELECT_MARGIN = 0.98  # the head must save at least 2% of the constant's validation iterations


def head_helps(net, val, data, w_const):
    """Whether the weight head lowers the validation instances' total
    iteration count (each at the grid weight nearest its choice) to at most
    SELECT_MARGIN times the calibrated constant's; returns (selected, head
    total, constant total)."""
    head = sum(grid_iterations(data, k, float(np.exp(net.log_weight(f, ns, nt)))) for f, ns, nt, _, k in val)
    const = sum(grid_iterations(data, k, w_const) for _, _, _, _, k in val)
    return head <= SELECT_MARGIN * const, head, const


def train_weight_head(net, train, val, epochs=200, lr=3e-3, lr_decay=0.99, seed=0, log=print):
    """Adam on the weight loss, head only; returns (best validation mse,
    history) and leaves the parameters of the best validation epoch set."""
    rng = np.random.default_rng(seed)
    best, best_params, hist = np.inf, net.get_params(), []
    for ep in range(epochs):
        tot = 0.0
        for i in rng.permutation(len(train)):
            f, ns, nt, t, _ = train[i]
            loss, _, _ = net.weight_loss_grad(f, ns, nt, t, False)
            net.adam_step(lr)
            tot += loss
        vl = float(np.mean([net.weight_loss_grad(f, ns, nt, t)[0] for f, ns, nt, t, _ in val]))
        hist.append((tot / len(train), vl))
        if vl < best:
            best, best_params = vl, net.get_params()
        if ep % 10 == 0 or ep == epochs - 1:
            log(f"head epoch {ep}: train mse {tot / len(train):.4f} val mse {vl:.4f} (best {best:.4f}) lr {lr:.2e}")
        lr *= lr_decay
    net.set_params(best_params)
    return best, hist


def evaluate_weights(net, count, n, w_const, tol=1e-4, storage=_lib.STORAGE_F32, seed0=HELD_OUT_SEED0, log=print):
    """Held-out n x n instances: iterations and time to `tol` from the zero
    start at the reference's weight 1.0, at the calibrated constant and at the
    head's per-instance weight (prediction time included)."""
    from dataclasses import replace
    rows = []
    for k in range(count):
        inst = instances.gmm_dirichlet(n, seed=seed0 + k)
        p = inst.problem
        dev = _lib.DeviceSolver(n, n, storage=storage)
        dev.upload(p.cost_matrix, p.source_distribution, p.target_distribution)
        t0 = time.time()
        w_net = net.predict_weight(inst.x, inst.y, p.source_distribution, p.target_distribution)
        t_pred = time.time() - t0
        row = {"seed": seed0 + k, "network_weight_value": w_net}
        for name, w, extra in (("default_weight", 1.0, 0.0), ("calibrated_weight", w_const, 0.0),
                               ("network_weight", w_net, t_pred)):
            params = replace(default_parameters(tol, record_iteration_stats=False), primal_importance=float(w))
            prof = SolveProfile()
            out = optimize(params, p, storage=storage, solver=dev, profile=prof)
            row[name] = {"iterations": out.iteration_count, "seconds": prof.solve_sec + extra,
                         "termination": out.termination_string, "weight": float(w)}
            log(f"held-out {k} {name} (weight {w:.3g}): {out.iteration_count} it {prof.solve_sec + extra:.2f}s "
                f"{out.termination_string}")
        dev.close()
        rows.append(row)
    summary = {}
    for name in ("default_weight", "calibrated_weight", "network_weight"):
        it = np.array([r[name]["iterations"] for r in rows], dtype=float)
        summary[name] = {"iterations_total": int(it.sum()), "seconds_total": float(sum(r[name]["seconds"] for r in rows)),
                         "optimal": int(sum(r[name]["termination"] == "OPTIMAL" for r in rows))}
    ratio = [r["calibrated_weight"]["iterations"] / max(1, r["network_weight"]["iterations"]) for r in rows]
    summary["network_vs_calibrated_iteration_ratio_median"] = float(np.median(ratio))
    summary["network_vs_calibrated_time_ratio_total"] = (summary["calibrated_weight"]["seconds_total"]
                                                         / summary["network_weight"]["seconds_total"])
    return rows, summary


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--grid", type=int, default=1792, help="weight-data instances at n, m in [--n-min, --n-max]")
    ap.add_argument("--grid-4096", type=int, default=256, help="weight-data instances at 4096 x 4096")
    ap.add_argument("--n-min", type=int, default=1024)
    ap.add_argument("--n-max", type=int, default=2048)
    ap.add_argument("--head-epochs", type=int, default=60)
    ap.add_argument("--head-lr", type=float, default=3e-3)
    ap.add_argument("--train", type=int, default=0, help="dual-network training instances (0: keep the duals in --out)")
    ap.add_argument("--dual-n-min", type=int, default=256)
    ap.add_argument("--dual-n-max", type=int, default=768)
    ap.add_argument("--data-tol", type=float, default=1e-5)
    ap.add_argument("--epochs", type=int, default=25)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eval", type=int, default=8, help="held-out instances")
    ap.add_argument("--eval-n", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=learned.DEFAULT_WEIGHTS)
    ap.add_argument("--init", default=learned.DEFAULT_WEIGHTS,
                    help="weights file whose dual network is kept when --train is 0")
    ap.add_argument("--data", default=None,
                    help="comma-separated npz files of weight data to use instead of generating it")
    ap.add_argument("--data-out", default=None, help="write the generated weight data here")
    ap.add_argument("--data-only", action="store_true", help="stop after generating (and writing) the weight data")
    ap.add_argument("--grid-seed0", type=int, default=WEIGHT_SEED0, help="first seed of the weight-data instances")
    ap.add_argument("--report", default=None)
    args = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)  # noqa: E731
    report = {}

    if args.train > 0:
        t0 = time.time()
        samples = generate_training_data(args.train, args.dual_n_min, args.dual_n_max, args.data_tol, args.seed,
                                         log=log)
        rng = np.random.default_rng(args.seed)
        order = rng.permutation(len(samples))
        nval = max(1, len(samples) // 10)
        val = [samples[i] for i in order[:nval]]
        trn = [samples[i] for i in order[nval:]]
        net = learned.WarmStartNet(seed=args.seed)
        hist = learned.train(net, trn, epochs=args.epochs, lr=args.lr, seed=args.seed, log=log)
        report["duals"] = {"train_instances": len(trn), "val_mse": learned.evaluate_loss(net, val),
                           "train_loss_history": hist, "sec": time.time() - t0}
        meta = {"train_instances": len(trn), "val_mse": report["duals"]["val_mse"], "epochs": args.epochs}
    else:
        net = learned.WarmStartNet(learned.load_params(args.init))
        net.set_params(np.concatenate([net.get_params()[:learned.NUM_DUAL_PARAMS], learned.init_head(args.seed)]))
        meta = {k: v for k, v in learned.load_meta(args.init).items()
                if k in ("train_instances", "val_mse", "epochs")}

    # 1. per-instance weight data
    t0 = time.time()
    parts = []
    if args.data:
        for path in args.data.split(","):
            with np.load(path, allow_pickle=False) as z:
                parts.append({k: z[k] for k in z.files})
            log(f"weight data: {len(parts[-1]['seed'])} instances from {path}")
    else:
        if args.grid > 0:
            parts.append(weight_grid_data(args.grid, args.n_min, args.n_max, args.grid_seed0, log=log))
        if args.grid_4096 > 0:
            parts.append(weight_grid_data(args.grid_4096, 4096, 4096, args.grid_seed0 + 50000, fixed_n=4096,
                                          batch_size=64, log=log))
    if not parts:
        raise SystemExit("no weight data: give --data or --grid / --grid-4096")
    data = {k: np.concatenate([q[k] for q in parts]) for k in ("seed", "n", "m", "iterations")}
    data["grid"] = parts[0]["grid"]
    if any(not np.array_equal(q["grid"], data["grid"]) for q in parts):
        raise SystemExit("weight data files were made with different grids")
    if args.data_out:
        np.savez(args.data_out, **data)
    report["weight_data_sec"] = time.time() - t0
    if args.data_only:
        return 0
    solves = int((data["iterations"] != 0).sum())
    # 2. calibrated constant over the same data
    w_const = calibrated_constant(data)
    log(f"calibrated constant weight: {w_const:g} over {len(data['seed'])} instances ({solves} grid solves)")
    # 3. the weight head
    samples = weight_samples(data)
    rng = np.random.default_rng(args.seed + 1)
    order = rng.permutation(len(samples))
    nval = max(1, len(samples) // 10)
    val = [samples[i] for i in order[:nval]]
    trn = [samples[i] for i in order[nval:]]
    tv = np.array([s_[3] for s_ in val])
    const_mse = float(np.mean((tv - np.log(w_const)) ** 2))
    t0 = time.time()
    best_val, hist = train_weight_head(net, trn, val, epochs=args.head_epochs, lr=args.head_lr, seed=args.seed,
                                       log=log)
    selected, it_head, it_const = head_helps(net, val, data, w_const)
    log(f"head: validation mse {best_val:.4f} against {const_mse:.4f} for the calibrated constant and "
        f"{float(np.var(tv)):.4f} for the mean target; validation iterations {it_head:.0f} (head) against "
        f"{it_const:.0f} (constant): head {'selected' if selected else 'not selected'}")
    meta.update({"initial_primal_weight": w_const, "weight_head": 1, "weight_head_selected": int(selected),
                 "weight_grid": data["grid"], "weight_train_instances": len(trn), "weight_val_instances": len(val),
                 "weight_val_mse": best_val, "weight_val_mse_constant": const_mse,
                 "weight_val_target_variance": float(np.var(tv)), "weight_val_iterations_head": it_head,
                 "weight_val_iterations_constant": it_const, "weight_grid_solves": solves})
    learned.save_params(args.out, net.get_params(), meta)
    report.update({"grid_solves": solves, "weight_train_instances": len(trn), "weight_val_instances": len(val),
                   "calibrated_constant": w_const, "head_val_mse": best_val, "constant_val_mse": const_mse,
                   "val_target_variance": float(np.var(tv)), "head_selected": bool(selected),
                   "val_iterations_head": it_head, "val_iterations_constant": it_const,
                   "head_train_sec": time.time() - t0, "head_history": hist})
    # 4. held-out evaluation
    if args.eval > 0:
        rows, summary = evaluate_weights(net, args.eval, args.eval_n, w_const, log=log)
        report.update({"eval_n": args.eval_n, "eval": rows, "summary": summary})
        log(json.dumps(summary))
    net.close()
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1, default=float)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
