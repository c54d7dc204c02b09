"""ctypes binding of libpdot_hip.so (declarations in include/pdot.h).

The HIP library is the only compute path: loading fails loudly when the
shared object is missing or was built for another ABI.  There is no CPU
fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpdot_hip.so")

PDOT_OK = 0
PDOT_E_ARG = -1
PDOT_E_HIP = -2
PDOT_E_STATE = -3
PDOT_E_NOMEM = -4

STORAGE_F64 = 0
STORAGE_F32 = 1
POLICY_ADAPTIVE = 0
POLICY_CONSTANT = 1
POINT_CURRENT = 0
POINT_AVERAGE = 1


class SolverState(ctypes.Structure):
    """pdot_solver_state (include/pdot.h)."""

    _fields_ = [
        ("step_size", ctypes.c_double),
        ("primal_weight", ctypes.c_double),
        ("cumulative_kkt_passes", ctypes.c_double),
        ("sum_weights", ctypes.c_double),
        ("total_number_iterations", ctypes.c_int64),
        ("avg_count", ctypes.c_int64),
        ("take_step_calls", ctypes.c_int64),
        ("error_call", ctypes.c_int64),
        ("last_interaction", ctypes.c_double),
        ("last_movement", ctypes.c_double),
        ("last_step_limit", ctypes.c_double),
        ("last_dx2", ctypes.c_double),
        ("last_dot", ctypes.c_double),
        ("last_dy2", ctypes.c_double),
        ("reduction_exponent", ctypes.c_double),
        ("growth_exponent", ctypes.c_double),
        ("step_policy", ctypes.c_int32),
        ("numerical_error", ctypes.c_int32),
        ("gzero", ctypes.c_int32),
        ("cur_slot", ctypes.c_int32),
        ("avg_slot", ctypes.c_int32),
        ("pending", ctypes.c_int32),
        ("spare0", ctypes.c_int32),
        ("spare1", ctypes.c_int32),
        ("step_budget", ctypes.c_int64),
        ("defer_ok", ctypes.c_int32),
        ("avg_pending", ctypes.c_int32),
        ("pend_nw", ctypes.c_double),
        ("pend_aw", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class PointStats(ctypes.Structure):
    """pdot_point_stats (include/pdot.h)."""

    _fields_ = [
        ("primal_objective", ctypes.c_double),
        ("x_inf", ctypes.c_double),
        ("x_sq", ctypes.c_double),
        ("rc_inf", ctypes.c_double),
        ("rc_sq", ctypes.c_double),
        ("p_inf", ctypes.c_double),
        ("p_sq", ctypes.c_double),
        ("dual_objective", ctypes.c_double),
        ("u_inf", ctypes.c_double),
        ("u_sq", ctypes.c_double),
        ("y_sq", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


_c_double_p = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ctx_p = ctypes.c_void_p

# name -> (restype, argtypes); the exported C-ABI surface
SIGNATURES = {
    "pdot_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int64, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32]),
    "pdot_destroy": (None, [_ctx_p]),
    "pdot_last_error": (ctypes.c_char_p, [_ctx_p]),
    "pdot_upload": (ctypes.c_int, [_ctx_p, _c_double_p, _c_double_p, _c_double_p]),
    "pdot_cached_ot_info": (ctypes.c_int, [_ctx_p, _c_double_p]),
    "pdot_upload_points": (ctypes.c_int, [_ctx_p, ctypes.c_int32, ctypes.c_int32, _c_double_p, _c_double_p, ctypes.c_int32, _c_double_p, _c_double_p]),
    "pdot_set_start": (ctypes.c_int, [_ctx_p, _c_double_p, _c_double_p, ctypes.c_int32, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    "pdot_run_steps": (ctypes.c_int, [_ctx_p, ctypes.c_int64, ctypes.POINTER(SolverState)]),
    "pdot_eval": (ctypes.c_int, [_ctx_p, ctypes.POINTER(PointStats), ctypes.POINTER(PointStats), ctypes.POINTER(SolverState)]),
    "pdot_discard": (ctypes.c_int, [_ctx_p]),
    "pdot_restart": (ctypes.c_int, [_ctx_p, ctypes.c_int32, ctypes.c_double]),
    "pdot_get_state": (ctypes.c_int, [_ctx_p, ctypes.POINTER(SolverState)]),
    "pdot_set_state": (ctypes.c_int, [_ctx_p, ctypes.POINTER(SolverState)]),
    "pdot_download": (ctypes.c_int, [_ctx_p, ctypes.c_int32, _c_double_p, _c_double_p, _c_double_p]),
    "pdot_time_steps": (ctypes.c_int, [_ctx_p, ctypes.c_int64, _c_double_p, _c_double_p]),
    "pdot_use_graphs": (ctypes.c_int, [_ctx_p, ctypes.c_int32]),
    "pdot_set_defer": (ctypes.c_int, [_ctx_p, ctypes.c_int32]),
    "pdot_set_pass_kernels": (ctypes.c_int, [_ctx_p, ctypes.c_int32]),
    "pdot_set_controller_fold": (ctypes.c_int, [_ctx_p, ctypes.c_int32]),
    "pdot_set_sparse": (ctypes.c_int, [_ctx_p, ctypes.c_int32]),
    "pdot_zero_share": (ctypes.c_int, [_ctx_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]),
    "pdot_set_stream_policy": (ctypes.c_int, [_ctx_p, ctypes.c_int32]),
    "pdot_set_launch": (ctypes.c_int, [_ctx_p, ctypes.c_int32, ctypes.c_int32]),
    "pdot_get_launch": (ctypes.c_int, [_ctx_p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
    "pdot_device_bytes": (ctypes.c_int64, [_ctx_p]),
    "pdot_copy_probe": (ctypes.c_int, [_ctx_p, ctypes.c_int64, ctypes.c_int32, _c_double_p]),
    # rounding and output
    "pdot_round": (ctypes.c_int, [_ctx_p, ctypes.c_int32, _c_double_p]),
    "pdot_round_matrix": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, _c_double_p, _c_double_p, _c_double_p, ctypes.c_char_p, ctypes.c_int64]),
    "pdot_round_device_noalloc": (ctypes.c_int, [ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_char_p, ctypes.c_int64]),
    "pdot_write_julia_vector": (ctypes.c_int, [ctypes.c_char_p, _c_double_p, ctypes.c_int64, ctypes.c_char_p, ctypes.c_int64]),
    "pdot_format_julia_float": (ctypes.c_int, [ctypes.c_double, ctypes.c_char_p, ctypes.c_int64]),
    # batches
    "pdot_batch_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int32, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64), ctypes.c_int32, ctypes.c_int32]),
    "pdot_batch_destroy": (None, [_ctx_p]),
    "pdot_batch_last_error": (ctypes.c_char_p, [_ctx_p]),
    "pdot_batch_upload": (ctypes.c_int, [_ctx_p, ctypes.c_int32, _c_double_p, _c_double_p, _c_double_p]),
    "pdot_batch_set_start": (ctypes.c_int, [_ctx_p, ctypes.c_int32, ctypes.c_double, ctypes.c_double, _c_double_p, _c_double_p]),
    "pdot_batch_set_warm_start": (ctypes.c_int, [_ctx_p, ctypes.c_int32, _c_double_p, _c_double_p]),
    "pdot_batch_set_active": (ctypes.c_int, [_ctx_p, ctypes.POINTER(ctypes.c_int32)]),
    "pdot_batch_run_steps": (ctypes.c_int, [_ctx_p, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(SolverState)]),
    "pdot_batch_eval": (ctypes.c_int, [_ctx_p, ctypes.POINTER(PointStats), ctypes.POINTER(PointStats), ctypes.POINTER(SolverState)]),
    "pdot_batch_restart": (ctypes.c_int, [_ctx_p, ctypes.POINTER(ctypes.c_int32), _c_double_p]),
    "pdot_batch_discard": (ctypes.c_int, [_ctx_p, ctypes.POINTER(ctypes.c_int32)]),
    "pdot_batch_get_states": (ctypes.c_int, [_ctx_p, ctypes.POINTER(SolverState)]),
    "pdot_batch_set_states": (ctypes.c_int, [_ctx_p, ctypes.POINTER(SolverState)]),
    "pdot_batch_download": (ctypes.c_int, [_ctx_p, ctypes.c_int32, ctypes.c_int32, _c_double_p, _c_double_p, _c_double_p]),
    "pdot_batch_time_steps": (ctypes.c_int, [_ctx_p, ctypes.c_int64, _c_double_p, _c_double_p]),
    "pdot_batch_use_graphs": (ctypes.c_int, [_ctx_p, ctypes.c_int32]),
    "pdot_batch_set_defer": (ctypes.c_int, [_ctx_p, ctypes.c_int32]),
    "pdot_batch_set_sparse": (ctypes.c_int, [_ctx_p, ctypes.c_int32]),
    "pdot_batch_set_launch": (ctypes.c_int, [_ctx_p, ctypes.c_int32, ctypes.c_int32]),
    "pdot_batch_get_launch": (ctypes.c_int, [_ctx_p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
    "pdot_batch_device_bytes": (ctypes.c_int64, [_ctx_p]),
    # learned warm start network

    "pdot_mlp_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int32]),
    "pdot_mlp_destroy": (None, [ctypes.c_void_p]),
    "pdot_mlp_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "pdot_mlp_num_params": (ctypes.c_int64, [ctypes.c_void_p]),
    "pdot_mlp_set_params": (ctypes.c_int, [ctypes.c_void_p, _fp]),
    "pdot_mlp_get_params": (ctypes.c_int, [ctypes.c_void_p, _fp]),
    "pdot_mlp_forward": (ctypes.c_int, [ctypes.c_void_p, _fp, ctypes.c_int64, ctypes.c_int64, _fp]),
    "pdot_mlp_loss_grad": (ctypes.c_int, [ctypes.c_void_p, _fp, ctypes.c_int64, ctypes.c_int64, _fp, _fp, _fp]),
    "pdot_mlp_adam_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float]),
    "pdot_mlp_predict_weight": (ctypes.c_int, [ctypes.c_void_p, _fp, ctypes.c_int64, ctypes.c_int64, _fp]),
    "pdot_mlp_weight_loss_grad": (ctypes.c_int, [ctypes.c_void_p, _fp, ctypes.c_int64, ctypes.c_int64, ctypes.c_float, ctypes.c_int32, _fp, _fp, _fp]),
    "pdot_mlp_gemm_probe": (ctypes.c_int, [ctypes.c_void_p, _fp, _fp, _fp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32]),
}

_lib = None


def load_library(path: str | None = None) -> ctypes.CDLL:
    """Load libpdot_hip.so; raises if it is absent (no fallback path exists)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"libpdot_hip.so not found at {p}: build it first (python pdot.jl_amd/build.py); "
            "the PDOT hot path has no CPU fallback")
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _ptr(a: np.ndarray | None):
    if a is None:
        return None
    return a.ctypes.data_as(_c_double_p)


class PdotError(RuntimeError):
    pass


class DeviceSolver:
    """One pdot_ctx: a problem resident in HBM plus its solver state."""

    def __init__(self, n: int, m: int, storage: int = STORAGE_F64, device: int = 0):
        self.lib = load_library()
        self.n, self.m, self.storage = int(n), int(m), int(storage)
        h = ctypes.c_void_p()
        rc = self.lib.pdot_create(ctypes.byref(h), self.n, self.m, self.storage, int(device))
        self.h = h
        if rc != PDOT_OK:
            msg = self.lib.pdot_last_error(h).decode() if h.value else f"code {rc}"
            if h.value:
                self.lib.pdot_destroy(h)
            self.h = ctypes.c_void_p()
            raise PdotError(f"pdot_create failed ({rc}): {msg}")

    def _check(self, rc: int, what: str) -> None:
        if rc != PDOT_OK:
            raise PdotError(f"{what} failed ({rc}): {self.lib.pdot_last_error(self.h).decode()}")

    def close(self) -> None:
        if self.h is not None and self.h.value:
            self.lib.pdot_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, C: np.ndarray, a: np.ndarray, b: np.ndarray) -> None:
        C = np.asfortranarray(C, dtype=np.float64)
        if C.shape != (self.n, self.m):
            raise ValueError(f"cost matrix shape {C.shape} != {(self.n, self.m)}")
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        if a.shape != (self.n,) or b.shape != (self.m,):
            raise ValueError("distribution lengths do not match the cost matrix")
        self._check(self.lib.pdot_upload(self.h, C.ctypes.data_as(_c_double_p), _ptr(a), _ptr(b)), "pdot_upload")

    def set_start(self, u0=None, v0=None, policy=POLICY_ADAPTIVE, reduction_exponent=2.0,
                  growth_exponent=1.0, step_size=1.0, primal_weight=1.0) -> None:
        if (u0 is None) != (v0 is None):
            raise ValueError("give both u0 and v0 or neither")
        if u0 is not None:
            u0 = np.ascontiguousarray(u0, dtype=np.float64)
            v0 = np.ascontiguousarray(v0, dtype=np.float64)
        self._check(self.lib.pdot_set_start(self.h, _ptr(u0), _ptr(v0), int(policy), float(reduction_exponent),
                                             float(growth_exponent), float(step_size), float(primal_weight)),
                    "pdot_set_start")

    def run_steps(self, count: int) -> SolverState:
        st = SolverState()
        self._check(self.lib.pdot_run_steps(self.h, int(count), ctypes.byref(st)), "pdot_run_steps")
        return st

    def eval(self):
        cur, avg, st = PointStats(), PointStats(), SolverState()
        self._check(self.lib.pdot_eval(self.h, ctypes.byref(cur), ctypes.byref(avg), ctypes.byref(st)), "pdot_eval")
        return cur, avg, st

    def discard(self) -> None:
        self._check(self.lib.pdot_discard(self.h), "pdot_discard")

    def restart(self, to_average: bool, primal_weight: float) -> None:
        self._check(self.lib.pdot_restart(self.h, int(bool(to_average)), float(primal_weight)), "pdot_restart")

    def get_state(self) -> SolverState:
        st = SolverState()
        self._check(self.lib.pdot_get_state(self.h, ctypes.byref(st)), "pdot_get_state")
        return st

    def set_state(self, st: SolverState) -> None:
        self._check(self.lib.pdot_set_state(self.h, ctypes.byref(st)), "pdot_set_state")

    def download(self, which: int = POINT_CURRENT, want_x: bool = True):
        X = np.empty((self.n, self.m), dtype=np.float64, order="F") if want_x else None
        u = np.empty(self.n, dtype=np.float64)
        v = np.empty(self.m, dtype=np.float64)
        xp = X.ctypes.data_as(_c_double_p) if X is not None else None
        self._check(self.lib.pdot_download(self.h, int(which), xp, _ptr(u), _ptr(v)), "pdot_download")
        return X, u, v

    def upload_points(self, problem) -> None:
        """Point-cloud problem (io.PointCloudProblem): the cost is formed on
        the device each pass and never stored."""
        from .io import COST_KINDS
        x, y = problem.source_points, problem.target_points
        if x.shape[0] != self.n or y.shape[0] != self.m:
            raise ValueError(f"point clouds {x.shape[0]} x {y.shape[0]} != {(self.n, self.m)}")
        self._check(self.lib.pdot_upload_points(self.h, COST_KINDS[problem.cost], x.shape[1], _ptr(x), _ptr(y),
                                                int(bool(problem.normalize)), _ptr(problem.source_distribution),
                                                _ptr(problem.target_distribution)), "pdot_upload_points")

    def cached_ot_info(self):
        """termination.cached_ot_info of the uploaded problem, from the device
        upload pass (src/termination.jl:94-101)."""
        from .termination import CachedOTInfo
        out = np.zeros(4)
        self._check(self.lib.pdot_cached_ot_info(self.h, _ptr(out)), "pdot_cached_ot_info")
        return CachedOTInfo(*(float(x) for x in out))

    def round(self, which: int = POINT_AVERAGE) -> np.ndarray:
        """Feasibility rounding of the resident current or average plan
        (src/utils_gpu.jl:296-318) computed on the device; FP64 n x m."""
        X = np.empty((self.n, self.m), dtype=np.float64, order="F")
        self._check(self.lib.pdot_round(self.h, int(which), X.ctypes.data_as(_c_double_p)), "pdot_round")
        return X

    def time_steps(self, count: int):
        ks, tt = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.pdot_time_steps(self.h, int(count), ctypes.byref(ks), ctypes.byref(tt)), "pdot_time_steps")
        return ks.value, tt.value

    def use_graphs(self, enable: bool) -> None:
        self._check(self.lib.pdot_use_graphs(self.h, int(bool(enable))), "pdot_use_graphs")

    def set_defer(self, enable: bool) -> None:
        """Deferred averaging on (default) or off; results are identical either way."""
        self._check(self.lib.pdot_set_defer(self.h, int(bool(enable))), "pdot_set_defer")

    def set_pass_kernels(self, enable: bool) -> None:
        self._check(self.lib.pdot_set_pass_kernels(self.h, int(bool(enable))), "pdot_set_pass_kernels")

    def set_controller_fold(self, mode: int) -> None:
        """The controller in the reduce kernel's last block (1), in a kernel of
        its own (0) or by size (-1, the default); results are identical."""
        self._check(self.lib.pdot_set_controller_fold(self.h, int(mode)), "pdot_set_controller_fold")

    def set_sparse(self, enable: bool) -> None:
        """Zero maps on (default: entries known to be +0.0 are neither read
        nor written) or off (every entry streamed); results are identical."""
        self._check(self.lib.pdot_set_sparse(self.h, int(bool(enable))), "pdot_set_sparse")

    def zero_share(self, which: int = POINT_CURRENT) -> float:
        """Share of the point's 16-byte column units recorded as +0.0 (skipped by passes)."""
        out = ctypes.c_double()
        self._check(self.lib.pdot_zero_share(self.h, int(which), ctypes.byref(out)), "pdot_zero_share")
        return out.value

    def set_stream_policy(self, mode: int) -> None:
        """Step-kernel streams with the nontemporal hint (0), without it (1) or
        by size (-1, the default); results are identical."""
        self._check(self.lib.pdot_set_stream_policy(self.h, int(mode)), "pdot_set_stream_policy")

    def set_launch(self, col_blocks_hint: int = 0, interleave: bool = True) -> None:
        self._check(self.lib.pdot_set_launch(self.h, int(col_blocks_hint), int(bool(interleave))), "pdot_set_launch")

    def get_launch(self) -> dict:
        vals = [ctypes.c_int32() for _ in range(4)]
        self._check(self.lib.pdot_get_launch(self.h, *[ctypes.byref(v) for v in vals]), "pdot_get_launch")
        return dict(zip(("rows_per_block", "col_blocks", "cols_per_block", "blocks"), (v.value for v in vals)))

    def device_bytes(self) -> int:
        return int(self.lib.pdot_device_bytes(self.h))

    def copy_probe(self, nbytes: int, reps: int = 10) -> float:
        """Mean milliseconds of a plain device copy of nbytes (2*nbytes of traffic)."""
        ms = ctypes.c_double()
        self._check(self.lib.pdot_copy_probe(self.h, int(nbytes), int(reps), ctypes.byref(ms)), "pdot_copy_probe")
        return ms.value


def _i32(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def _i64(a):
    a = np.ascontiguousarray(a, dtype=np.int64)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


class BatchSolver:
    """One pdot_batch: many problems resident in HBM, advanced together."""

    def __init__(self, ns, ms, storage: int = STORAGE_F32, device: int = 0):
        self.lib = load_library()
        self.ns = [int(x) for x in ns]
        self.ms = [int(x) for x in ms]
        self.count = len(self.ns)
        if self.count == 0 or len(self.ms) != self.count:
            raise ValueError("need the same, non-zero number of row and column sizes")
        self.storage = int(storage)
        self.sparse = True
        na, npp = _i64(self.ns)
        ma, mp = _i64(self.ms)
        h = ctypes.c_void_p()
        rc = self.lib.pdot_batch_create(ctypes.byref(h), self.count, npp, mp, self.storage, int(device))
        self.h = h
        if rc != PDOT_OK:
            msg = self.lib.pdot_batch_last_error(h).decode() if h.value else f"code {rc}"
            if h.value:
                self.lib.pdot_batch_destroy(h)
            self.h = ctypes.c_void_p()
            raise PdotError(f"pdot_batch_create failed ({rc}): {msg}")

    def _check(self, rc: int, what: str) -> None:
        if rc != PDOT_OK:
            raise PdotError(f"{what} failed ({rc}): {self.lib.pdot_batch_last_error(self.h).decode()}")

    def close(self) -> None:
        if self.h is not None and self.h.value:
            self.lib.pdot_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, k: int, C, a, b) -> None:
        C = np.asfortranarray(C, dtype=np.float64)
        if C.shape != (self.ns[k], self.ms[k]):
            raise ValueError(f"instance {k}: cost matrix shape {C.shape} != {(self.ns[k], self.ms[k])}")
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        self._check(self.lib.pdot_batch_upload(self.h, int(k), C.ctypes.data_as(_c_double_p), _ptr(a), _ptr(b)),
                    "pdot_batch_upload")

    def set_start(self, policy, reduction_exponent, growth_exponent, step_size, primal_weight) -> None:
        eta = np.ascontiguousarray(np.broadcast_to(np.asarray(step_size, dtype=np.float64), (self.count,)))
        om = np.ascontiguousarray(np.broadcast_to(np.asarray(primal_weight, dtype=np.float64), (self.count,)))
        self._check(self.lib.pdot_batch_set_start(self.h, int(policy), float(reduction_exponent),
                                                  float(growth_exponent), _ptr(eta), _ptr(om)), "pdot_batch_set_start")

    def set_warm_start(self, k: int, u0, v0) -> None:
        u0 = np.ascontiguousarray(u0, dtype=np.float64)
        v0 = np.ascontiguousarray(v0, dtype=np.float64)
        self._check(self.lib.pdot_batch_set_warm_start(self.h, int(k), _ptr(u0), _ptr(v0)), "pdot_batch_set_warm_start")

    def set_active(self, active) -> None:
        a, p = _i32(active)
        self._check(self.lib.pdot_batch_set_active(self.h, p), "pdot_batch_set_active")

    def run_steps(self, counts):
        c, p = _i64(counts)
        st = (SolverState * self.count)()
        self._check(self.lib.pdot_batch_run_steps(self.h, p, st), "pdot_batch_run_steps")
        return st

    def eval(self):
        cur, avg, st = (PointStats * self.count)(), (PointStats * self.count)(), (SolverState * self.count)()
        self._check(self.lib.pdot_batch_eval(self.h, cur, avg, st), "pdot_batch_eval")
        return cur, avg, st

    def restart(self, actions, weights) -> None:
        a, p = _i32(actions)
        w = np.ascontiguousarray(weights, dtype=np.float64)
        self._check(self.lib.pdot_batch_restart(self.h, p, _ptr(w)), "pdot_batch_restart")

    def discard(self, which) -> None:
        a, p = _i32(which)
        self._check(self.lib.pdot_batch_discard(self.h, p), "pdot_batch_discard")

    def get_states(self):
        st = (SolverState * self.count)()
        self._check(self.lib.pdot_batch_get_states(self.h, st), "pdot_batch_get_states")
        return st

    def download(self, k: int, which: int = POINT_CURRENT, want_x: bool = True):
        n, m = self.ns[k], self.ms[k]
        X = np.empty((n, m), dtype=np.float64, order="F") if want_x else None
        u, v = np.empty(n), np.empty(m)
        xp = X.ctypes.data_as(_c_double_p) if X is not None else None
        self._check(self.lib.pdot_batch_download(self.h, int(k), int(which), xp, _ptr(u), _ptr(v)), "pdot_batch_download")
        return X, u, v

    def time_steps(self, count: int):
        ks, tt = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.pdot_batch_time_steps(self.h, int(count), ctypes.byref(ks), ctypes.byref(tt)),
                    "pdot_batch_time_steps")
        return ks.value, tt.value

    def use_graphs(self, enable: bool) -> None:
        self._check(self.lib.pdot_batch_use_graphs(self.h, int(bool(enable))), "pdot_batch_use_graphs")

    def set_defer(self, enable: bool) -> None:
        """Deferred averaging on (default) or off; results are identical either way."""
        self._check(self.lib.pdot_batch_set_defer(self.h, int(bool(enable))), "pdot_batch_set_defer")

    def set_sparse(self, enable: bool) -> None:
        """Zero maps on (default) or off for every instance; results are identical."""
        self._check(self.lib.pdot_batch_set_sparse(self.h, int(bool(enable))), "pdot_batch_set_sparse")
        self.sparse = bool(enable)

    def set_launch(self, cols_per_tile: int, grid_mult: int = 1) -> None:
        self._check(self.lib.pdot_batch_set_launch(self.h, int(cols_per_tile), int(grid_mult)), "pdot_batch_set_launch")

    def get_launch(self) -> dict:
        vals = [ctypes.c_int32() for _ in range(4)]
        self._check(self.lib.pdot_batch_get_launch(self.h, *[ctypes.byref(v) for v in vals]), "pdot_batch_get_launch")
        return dict(zip(("rows_per_tile", "cols_per_tile", "tiles", "step_grid"), (v.value for v in vals)))

    def device_bytes(self) -> int:
        return int(self.lib.pdot_batch_device_bytes(self.h))
