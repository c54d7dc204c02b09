"""Known answer of the third evaluation for a point-cloud problem, computed
on the host in column blocks (no n x m matrix held at once).

From the zero start, pass 1 leaves X = 0 with u = a, v = b and the step
size 1 -> 1.5; iteration 2 restarts to the average with the primal weight
1; pass 2 then gives X2 = max(-(1.5 ((C - a_i) - b_j)), 0), which the
iteration-3 evaluation reports as the average: ||X2||_F and <C, X2>
(the oracle's own known-answer test pins this sequence on small problems).
"""
import numpy as np


def third_evaluation(pts, block: int = 512):
    """(||X2||_F, <C, X2>) for a PointCloudProblem with normalize=False."""
    assert not pts.normalize
    x, y = pts.source_points, pts.target_points
    a, b = pts.source_distribution, pts.target_distribution
    f = np.square if pts.cost == "sqeuclid" else np.abs
    ss = obj = 0.0
    for j0 in range(0, y.shape[0], block):
        j1 = min(y.shape[0], j0 + block)
        C = np.zeros((x.shape[0], j1 - j0))
        for d in range(x.shape[1]):
            C += f(y[None, j0:j1, d] - x[:, d, None])
        X2 = np.maximum(-(1.5 * ((C - a[:, None]) - b[None, j0:j1])), 0.0)
        ss += float(np.sum(X2 * X2))
        obj += float(np.sum(C * X2))
    return float(np.sqrt(ss)), obj
