"""Batched solves: many independent problems of uneven size advanced together
on one GPU (pdot_batch in libpdot_hip.so).  Every instance follows exactly
the single-problem semantics of `optimize` (src/solver_gpu.jl:413-720): its
own step size, primal weight, restart scheme and termination.  All instances
share the evaluation schedule; an instance that terminates drops out of the
device work lists, and an instance halted by a numerical error keeps its
state frozen until the next evaluation, which is then reported at the
iteration the reference would have evaluated it (error_call + 1).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from .io import OptimalTransportProblem
from .solve_log import RestartChoice, termination_reason_to_string
from .solver import (InstanceController, PdhgParameters, SaddlePointOutput, initial_step_and_weight,
                     next_evaluation_iteration)
from .termination import cached_ot_info


@dataclass
class BatchProfile:
    upload_sec: float = 0.0
    solve_sec: float = 0.0
    output_sec: float = 0.0
    evaluations: int = 0
    passes_launched: int = 0
    finish_sec: List[float] = field(default_factory=list)


def autotune(batch: _lib.BatchSolver, candidates=((32, 1), (64, 1), (128, 1), (256, 1), (64, 2), (128, 2)),
             steps: int = 8, restore_start=None) -> dict:
    """Pick the step kernel's tile width and grid multiple by timing each
    candidate on the batch itself (a few passes from the zero start, with the
    zero maps off so that every element is read and written, as in the dense
    phase of a solve); leaves the fastest applied.  restore_start(batch)
    re-initialises the iterate afterwards (default: zero start with step 1,
    weight 1)."""
    sparse = batch.sparse
    batch.set_sparse(False)
    results = []
    for cols, mult in candidates:
        batch.set_launch(cols, mult)
        batch.set_start(_lib.POLICY_ADAPTIVE, 2.0, 1.0, 1.0, 1.0)
        batch.time_steps(2)
        ks, tt = batch.time_steps(steps)
        results.append({"cols_per_tile": cols, "grid_mult": mult, "ms_per_step": tt / steps,
                        "step_kernel_ms": ks / steps, **batch.get_launch()})
    best = min(results, key=lambda r: r["ms_per_step"])
    batch.set_launch(best["cols_per_tile"], best["grid_mult"])
    batch.set_sparse(sparse)
    if restore_start is not None:
        restore_start(batch)
    return {"best": best, "trials": results}


def optimize_batch(params: PdhgParameters, problems: Sequence[OptimalTransportProblem], *,
                   storage: int = _lib.STORAGE_F32, device: int = 0, warm_starts=None,
                   batch: Optional[_lib.BatchSolver] = None, use_graphs: bool = True,
                   profile: Optional[BatchProfile] = None) -> List[SaddlePointOutput]:
    """optimize() for every problem, advanced together.  warm_starts: optional
    list of (u0, v0) or None per problem.  batch: reuse a BatchSolver whose
    problems are already uploaded.  params may also be a list with one entry
    per problem (for example per-instance initial primal weights); the entries
    must share the evaluation frequency, and their iteration limits must equal
    the first entry's or be multiples of the frequency, so that every
    instance's evaluations fall on the common schedule."""
    count = len(problems)
    plist = list(params) if isinstance(params, (list, tuple)) else [params] * count
    if len(plist) != count:
        raise ValueError("one PdhgParameters per problem expected")
    params = plist[0]
    freq = params.termination_evaluation_frequency
    for p in plist:
        lim = p.termination_criteria.iteration_limit
        if p.termination_evaluation_frequency != freq:
            raise ValueError("every instance needs the same termination_evaluation_frequency")
        if p.step_size_policy_params != params.step_size_policy_params:
            raise ValueError("every instance needs the same step-size policy")
        if lim != params.termination_criteria.iteration_limit and not (lim > 10 and lim % freq == 0):
            raise ValueError("per-instance iteration limits must be multiples of the evaluation frequency")
    caches = [cached_ot_info(p) for p in problems]
    t_up = time.time()
    if batch is None:
        batch = _lib.BatchSolver([p.num_row for p in problems], [p.num_col for p in problems], storage=storage,
                                 device=device)
        for k, p in enumerate(problems):
            batch.upload(k, p.cost_matrix, p.source_distribution, p.target_distribution)
    upload_sec = time.time() - t_up
    batch.use_graphs(use_graphs)
    init = [initial_step_and_weight(plist[k], p.num_row, p.num_col, caches[k]) for k, p in enumerate(problems)]
    policy, re, ge = init[0][0], init[0][1], init[0][2]
    batch.set_start(policy, re, ge, [x[3] for x in init], [x[4] for x in init])
    if warm_starts is not None:
        for k, ws in enumerate(warm_starts):
            if ws is not None:
                batch.set_warm_start(k, ws[0], ws[1])
    ctls = [InstanceController(plist[k], caches[k]) for k in range(count)]
    outputs: List[Optional[SaddlePointOutput]] = [None] * count
    finish = [0.0] * count
    active = np.ones(count, dtype=np.int32)
    iteration_limit = params.termination_criteria.iteration_limit
    start_time = time.time()
    time_basic = 0.0
    passes = 0
    iteration = 1
    while active.any():
        t0 = time.time()
        cur, avg, st = batch.eval()
        time_basic += time.time() - t0
        elapsed = time.time() - start_time
        actions = np.zeros(count, dtype=np.int32)
        weights = np.zeros(count)
        done_now = np.zeros(count, dtype=np.int32)
        for k in np.flatnonzero(active):
            s = st[k]
            it_k = int(s.error_call) + 1 if s.numerical_error else iteration
            reason, stats = ctls[k].evaluate(cur[k], avg[k], s, it_k, elapsed, time_basic)
            if reason is not None:
                done_now[k] = 1
                outputs[k] = (reason, it_k, int(s.avg_count))
                finish[k] = elapsed
                continue
            decision, w = ctls[k].restart(cur[k], avg[k], s, it_k, stats)
            if decision.choice != RestartChoice.RESTART_CHOICE_NO_RESTART:
                actions[k] = 2 if decision.to_average else 1
                weights[k] = w
        if done_now.any():
            batch.discard(done_now)  # finished instances keep their output point resident
            active[done_now == 1] = 0
            batch.set_active(active)
            if not active.any():
                break
        if actions.any():
            batch.restart(actions, weights)
        nxt = next_evaluation_iteration(iteration, freq, iteration_limit)
        counts = np.where(actions != 0, nxt - iteration, nxt - iteration - 1).astype(np.int64)
        counts[active == 0] = 0
        t0 = time.time()
        batch.run_steps(counts)
        time_basic += time.time() - t0
        passes += int(counts.max()) if counts.size else 0
        iteration = nxt
    solve_sec = time.time() - start_time
    for k in range(count):  # outputs copied once every instance has terminated
        reason, it_k, avg_count = outputs[k]
        which = _lib.POINT_CURRENT if avg_count == 0 else _lib.POINT_AVERAGE
        X, u, v = batch.download(int(k), which)
        outputs[k] = SaddlePointOutput(X, u, v, reason, termination_reason_to_string(reason), it_k - 1,
                                       ctls[k].iteration_stats)
    if profile is not None:
        profile.upload_sec = upload_sec
        profile.solve_sec = solve_sec
        profile.output_sec = time.time() - start_time - solve_sec
        profile.evaluations = max(c.evaluations for c in ctls)
        profile.passes_launched = passes
        profile.finish_sec = finish
    return outputs
