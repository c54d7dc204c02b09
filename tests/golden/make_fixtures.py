"""Writes tests/golden/lp_optima.json: exact optimal values of small transport
LPs from independent solvers, used to pin the oracle and the GPU results.

  * scipy.optimize.linprog (HiGHS) on the transport LP min <C,X>,
    X 1 = a, X^T 1 = b, X >= 0, for instances from pdot_amd.instances;
  * scipy.optimize.linear_sum_assignment for n = m with uniform masses 1/n
    (Birkhoff: the LP optimum is an assignment, value = cost / n).

Run from the repository root:  python tests/golden/make_fixtures.py
"""
import json
import os
import sys

import numpy as np
import scipy
import scipy.sparse as sp
from scipy.optimize import linear_sum_assignment, linprog

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "pdot.jl_amd"))
from pdot_amd import instances  # noqa: E402

CASES = [
    # (generator, n, m, seed)
    ("uniform_sqeuclid", 300, 257, 11),
    ("uniform_sqeuclid", 1000, 1000, 1000),
    ("uniform_sqeuclid", 64, 700, 5),
    ("uniform_sqeuclid", 200, 150, 9),
    ("uniform_sqeuclid", 500, 800, 9),
    ("uniform_sqeuclid", 100, 80, 7),
    ("uniform_sqeuclid", 1, 50, 3),
    ("uniform_sqeuclid", 40, 1, 4),
    ("gaussian_shift", 400, 400, 2000),
    ("l1_mixture_3d", 300, 300, 3000),
    ("gmm_dirichlet", 256, 256, 5000),
]


def make_problem(gen, n, m, seed):
    out = instances.WORKLOADS[gen](n, m, seed=seed)
    return out.problem if hasattr(out, "problem") else out


def lp_optimum(prob):
    C, a, b = prob.cost_matrix, prob.source_distribution, prob.target_distribution
    n, m = C.shape
    # x is X in row-major order: x[i*m + j] = X[i, j]
    A = sp.vstack([sp.kron(sp.eye(n), np.ones((1, m))), sp.kron(np.ones((1, n)), sp.eye(m))]).tocsr()
    res = linprog(C.ravel(order="C"), A_eq=A, b_eq=np.concatenate([a, b]), bounds=(0, None), method="highs")
    assert res.status == 0, res.message
    return float(res.fun)


def assignment_case(n, seed):
    rng = np.random.default_rng(seed)
    x, y = rng.random((n, 2)), rng.random((n, 2))
    C = ((x[:, None, :] - y[None, :, :]) ** 2).sum(-1)
    r, c = linear_sum_assignment(C)
    return C, float(C[r, c].sum() / n)


def main():
    out = {"scipy": scipy.__version__, "numpy": np.__version__, "lp": [], "assignment": []}
    for gen, n, m, seed in CASES:
        prob = make_problem(gen, n, m, seed)
        val = lp_optimum(prob)
        out["lp"].append({"generator": gen, "n": n, "m": m, "seed": seed, "optimum": val,
                          "cost_checksum": float(prob.cost_matrix.sum())})
        print(gen, n, m, seed, val, flush=True)
    for n, seed in [(50, 1), (200, 2)]:
        C, val = assignment_case(n, seed)
        out["assignment"].append({"n": n, "seed": seed, "optimum": val, "cost_checksum": float(C.sum())})
        print("assignment", n, val)
    with open(os.path.join(HERE, "lp_optima.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
