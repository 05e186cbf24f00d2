"""Python bindings (ctypes) for the sankhya_gpu C ABI and its workload library.

The structures mirror ``include/sankhya_gpu.h`` field for field; that header
mirrors the reference C API (``c/include/sankhya.h:17-161``) and GPU ABI
(``cuda/include/sankhya_cuda.h``).  Nothing here computes: every solve goes
through ``libsankhya_gpu.so`` (HIP, gfx950).  Loading fails loudly when the
library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
GPU_LIB = HERE / "libsankhya_gpu.so"
WORKLOADS_LIB = HERE / "libsk_workloads.so"

c_int_p = C.POINTER(C.c_int)
c_double_p = C.POINTER(C.c_double)

# sk_status / sk_result values (c/include/sankhya.h:17-38)
SK_OK, SK_ERR_ARG, SK_ERR_MEMORY, SK_ERR_STRUCTURE = 0, 1, 2, 3
SK_ERR_IO, SK_ERR_SINGULAR, SK_ERR_UNSUPPORTED, SK_ERR_NUMERIC = 4, 5, 6, 7
RESULT_NAMES = ["optimal", "infeasible", "unbounded", "iteration_limit",
                "time_limit", "numeric_failure", "not_run", "gap_limit"]
SK_RESULT_OPTIMAL, SK_RESULT_INFEASIBLE, SK_RESULT_UNBOUNDED = 0, 1, 2
SK_RESULT_ITERATION_LIMIT, SK_RESULT_TIME_LIMIT, SK_RESULT_NUMERIC_FAILURE = 3, 4, 5
SK_RESULT_NOT_RUN = 6
SK_LP_AUTO, SK_LP_SIMPLEX, SK_LP_FIRST_ORDER = 0, 1, 2
SK_INFINITY = 1.0e30
PREC_FP64, PREC_MIXED = 0, 1
RESTART_KKT, RESTART_NDG = 0, 1


class SkCsc(C.Structure):
    _fields_ = [("nrow", C.c_int), ("ncol", C.c_int), ("p", c_int_p),
                ("i", c_int_p), ("x", c_double_p), ("nzmax", C.c_int)]


class SkModel(C.Structure):
    _fields_ = [("nrow", C.c_int), ("ncol", C.c_int), ("A", SkCsc),
                ("Q", C.POINTER(SkCsc)), ("c", c_double_p), ("rlow", c_double_p),
                ("rupp", c_double_p), ("clow", c_double_p), ("cupp", c_double_p),
                ("vartype", C.POINTER(C.c_ubyte)), ("objshift", C.c_double),
                ("objsense", C.c_double), ("rowname", C.POINTER(C.c_char_p)),
                ("colname", C.POINTER(C.c_char_p)), ("name", C.c_char_p)]


class SkSolution(C.Structure):
    _fields_ = [("result", C.c_int), ("objective", C.c_double), ("x", c_double_p),
                ("y", c_double_p), ("rc", c_double_p), ("rowact", c_double_p),
                ("ncol", C.c_int), ("nrow", C.c_int), ("iterations", C.c_longlong),
                ("nodes", C.c_longlong), ("solve_seconds", C.c_double),
                ("primal_infeasibility", C.c_double), ("dual_infeasibility", C.c_double),
                ("complementarity", C.c_double), ("mip_gap", C.c_double),
                ("dual_bound", C.c_double)]


class SkOptions(C.Structure):
    _fields_ = [("primal_tol", C.c_double), ("dual_tol", C.c_double),
                ("pivot_tol", C.c_double), ("time_limit", C.c_double),
                ("iteration_limit", C.c_longlong), ("node_limit", C.c_longlong),
                ("presolve", C.c_int), ("scaling", C.c_int),
                ("refactor_interval", C.c_int), ("verbosity", C.c_int),
                ("threads", C.c_int), ("mip_gap_abs", C.c_double),
                ("mip_gap_rel", C.c_double), ("mip_cuts", C.c_int),
                ("mip_heuristics", C.c_int), ("dual_simplex", C.c_int),
                ("lp_engine", C.c_int), ("random_seed", C.c_ulonglong)]


class GpuCsr(C.Structure):
    _fields_ = [("rows", C.c_int), ("cols", C.c_int), ("nnz", C.c_int),
                ("device_row_offsets", C.c_void_p), ("device_column_indices", C.c_void_p),
                ("device_values", C.c_void_p)]


class LpSettings(C.Structure):
    _fields_ = [("max_iterations", C.c_int), ("check_every", C.c_int),
                ("tau", C.c_double), ("sigma", C.c_double), ("theta", C.c_double),
                ("tolerance", C.c_double), ("time_limit", C.c_double)]


class LpResult(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_int), ("objective", C.c_double),
                ("maximum_row_violation", C.c_double), ("maximum_step", C.c_double),
                ("maximum_kkt_residual", C.c_double), ("maximum_dual_residual", C.c_double),
                ("maximum_complementarity", C.c_double)]


class PdlpParams(C.Structure):
    _fields_ = [("eps_optimal", C.c_double), ("max_iterations", C.c_int),
                ("restart_period", C.c_int), ("time_limit_seconds", C.c_double),
                ("power_iterations", C.c_int), ("adaptive_step", C.c_int),
                ("eta_min_factor", C.c_double), ("eta_max_factor", C.c_double),
                ("precision", C.c_int), ("tail_switch", C.c_double),
                ("tail_stall_windows", C.c_int), ("scaling", C.c_int),
                ("ruiz_iterations", C.c_int), ("ruiz_tol", C.c_double),
                ("use_graph", C.c_int), ("verbosity", C.c_int), ("primal_weight", C.c_double),
                ("unfused", C.c_int), ("refine", C.c_int), ("cancel", C.POINTER(C.c_int)),
                ("restart_rule", C.c_int), ("autotune", C.c_int), ("tile_nnz", C.c_int),
                ("spmv_grid", C.c_int), ("stream_loads", C.c_int), ("probe_grid", C.c_int),
                ("xcache_bytes", C.c_int), ("pad", C.c_int)]


class PdlpStats(C.Structure):
    _fields_ = [("converged", C.c_int), ("cancelled", C.c_int), ("iterations", C.c_int),
                ("restarts", C.c_int), ("primal_objective", C.c_double),
                ("dual_objective", C.c_double), ("relative_primal_residual", C.c_double),
                ("relative_dual_residual", C.c_double), ("relative_gap", C.c_double),
                ("matrix_norm", C.c_double), ("seconds", C.c_double),
                ("rejected_steps", C.c_int), ("final_step_size", C.c_double),
                ("step_size_ratio", C.c_double), ("host_syncs", C.c_int),
                ("fp32_iterations", C.c_int), ("fp64_iterations", C.c_int),
                ("kernel_launches", C.c_longlong), ("loop_seconds", C.c_double),
                ("setup_seconds", C.c_double), ("primal_weight", C.c_double),
                ("reanchors", C.c_int), ("iterations_to_merit", C.c_int * 4),
                ("seconds_to_merit", C.c_double * 4), ("restart_merit", C.c_double * 2),
                ("tile_nnz", C.c_int), ("spmv_grid", C.c_int), ("stream_loads", C.c_int), ("tuned", C.c_int),
                ("tune_seconds", C.c_double), ("probe_grid", C.c_int), ("xcache_bytes", C.c_int)]

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_}
        d["iterations_to_merit"] = list(self.iterations_to_merit)
        d["seconds_to_merit"] = list(self.seconds_to_merit)
        d["restart_merit"] = list(self.restart_merit)
        return d


class GpuOptions(C.Structure):
    _fields_ = [("pdlp", PdlpParams), ("x0", c_double_p), ("y0", c_double_p)]


class BatchOptions(C.Structure):
    _fields_ = [("pdlp", PdlpParams), ("block_threads", C.c_int), ("use_graph", C.c_int)]


class BatchResult(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_int), ("restarts", C.c_int),
                ("fp32_iterations", C.c_int), ("relative_primal_residual", C.c_double), ("relative_dual_residual", C.c_double),
                ("relative_gap", C.c_double), ("objective", C.c_double)]


class BatchStats(C.Structure):
    _fields_ = [("setup_seconds", C.c_double), ("device_seconds", C.c_double),
                ("seconds", C.c_double), ("block_threads", C.c_int), ("grid", C.c_int),
                ("lds_bytes", C.c_int), ("converged", C.c_int), ("total_iterations", C.c_longlong),
                ("tune_ms", C.c_double * 3), ("fp32_iterations", C.c_longlong), ("handoffs", C.c_int),
                ("grid_fp64", C.c_int), ("lds_bytes_fp64", C.c_int), ("pad", C.c_int)]

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_}
        d["tune_ms"] = list(self.tune_ms)
        return d


def default_pdlp_params(**overrides) -> PdlpParams:
    """Defaults of PdlpParams (src/cuda/GpuPdlp.hpp:19-67) except eps 1e-6;
    the same values the library's sk_gpu_pdlp_params_default sets."""
    p = PdlpParams(eps_optimal=1e-6, max_iterations=500000, restart_period=64,
                   time_limit_seconds=0.0, power_iterations=30, adaptive_step=1,
                   eta_min_factor=1e-6, eta_max_factor=1e4, precision=PREC_FP64,
                   tail_switch=1e-5, tail_stall_windows=32, scaling=1,
                   ruiz_iterations=100, ruiz_tol=1e-2, use_graph=0, verbosity=0, refine=1,
                   restart_rule=RESTART_KKT)
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None else None


def dptr(a):
    return _ptr(a, C.c_double)


def iptr(a):
    return _ptr(a, C.c_int)


# ---------------------------------------------------------------- loading

_gpu = None
_work = None


def _declare(lib, name, restype, argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = argtypes
    return f


def load_workloads():
    global _work
    if _work is None:
        if not WORKLOADS_LIB.exists():
            raise RuntimeError(f"{WORKLOADS_LIB} is not built; run __graft_entry__.build()")
        lib = C.CDLL(str(WORKLOADS_LIB))
        M = C.POINTER(SkModel)
        _declare(lib, "skw_model_free", None, [M])
        _declare(lib, "skw_bounded_lp", C.c_int, [M, C.c_int, C.c_int, C.c_int, C.c_ulonglong])
        _declare(lib, "skw_zipf_lp", C.c_int, [M, C.c_int, C.c_int, C.c_double, C.c_double,
                                               C.c_int, C.c_double, C.c_ulonglong, c_double_p])
        _declare(lib, "skw_sparse_qp", C.c_int, [M, C.c_int, C.c_int, C.c_int, C.c_int,
                                                 C.c_int, C.c_ulonglong])
        _declare(lib, "skw_batch_sizes", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_ulonglong, c_int_p])
        _declare(lib, "skw_parametric_family", C.c_int,
                 [M, C.c_int, C.c_int, C.c_int, C.c_int, C.c_ulonglong,
                  c_double_p, c_double_p, c_double_p, c_double_p])
        _declare(lib, "skw_parametric_instance", None,
                 [C.c_int, C.c_int, C.c_int, c_double_p, c_double_p, c_double_p, c_double_p,
                  c_double_p, c_double_p, c_double_p, c_double_p])
        _declare(lib, "skw_normal_matrix", None, [C.c_int, C.c_int, C.c_ulonglong, c_double_p])
        _work = lib
    return _work


def load_gpu():
    """Load libsankhya_gpu.so.  Raises if it is missing: there is no fallback."""
    global _gpu
    if _gpu is None:
        if not GPU_LIB.exists():
            raise RuntimeError(f"{GPU_LIB} is not built; run __graft_entry__.build()")
        lib = C.CDLL(str(GPU_LIB))
        M = C.POINTER(SkModel)
        S = C.POINTER(SkSolution)
        H = C.POINTER(GpuCsr)
        _declare(lib, "sk_gpu_last_error", C.c_char_p, [])
        _declare(lib, "sk_gpu_csr_create", C.c_int, [H, C.c_int, C.c_int, C.c_int, c_int_p, c_int_p, c_double_p])
        _declare(lib, "sk_gpu_csr_destroy", None, [H])
        _declare(lib, "sk_gpu_spmv_device_f64", C.c_int, [H, C.c_void_p, C.c_void_p])
        _declare(lib, "sk_gpu_spmv_transpose_device_f64", C.c_int, [H, C.c_void_p, C.c_void_p])
        _declare(lib, "sk_gpu_axpy_device_f64", C.c_int, [C.c_int, C.c_double, C.c_void_p, C.c_void_p])
        _declare(lib, "sk_gpu_synchronize", C.c_int, [])
        _declare(lib, "sk_gpu_relative_kkt_device_noalloc", C.c_int,
                 [H] + [C.c_void_p] * 8)
        _declare(lib, "sk_gpu_spmv_f64", C.c_int, [C.c_int, C.c_int, C.c_int, c_int_p, c_int_p,
                                                   c_double_p, c_double_p, c_double_p])
        _declare(lib, "sk_gpu_axpy_f64", C.c_int, [C.c_int, C.c_double, c_double_p, c_double_p])
        _declare(lib, "sk_gpu_spmv_f32", C.c_int, [C.c_int, C.c_int, C.c_int, c_int_p, c_int_p, c_double_p,
                                                   C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)])
        lp_args = [H, c_double_p, c_double_p, c_double_p, c_double_p, c_double_p, LpSettings,
                   c_double_p, C.POINTER(LpResult)]
        _declare(lib, "sk_gpu_lp_pdhg", C.c_int, lp_args)
        _declare(lib, "sk_gpu_diagonal_qp_pdhg", C.c_int,
                 [H, c_double_p, c_double_p, c_double_p, c_double_p, c_double_p, c_double_p,
                  LpSettings, c_double_p, C.POINTER(LpResult)])
        _declare(lib, "sk_gpu_sparse_qp_pdhg", C.c_int,
                 [H, H, c_double_p, c_double_p, c_double_p, c_double_p, c_double_p,
                  LpSettings, c_double_p, C.POINTER(LpResult)])
        _declare(lib, "sk_gpu_qp_pdhg_preconditioned", C.c_int,
                 [H, H, c_double_p, c_double_p, c_double_p, c_double_p, c_double_p, c_double_p,
                  c_double_p, c_double_p, LpSettings, c_double_p, C.POINTER(LpResult)])
        _declare(lib, "sk_gpu_qp_pdhg_preconditioned_with_dual", C.c_int,
                 [H, H, c_double_p, c_double_p, c_double_p, c_double_p, c_double_p, c_double_p,
                  c_double_p, c_double_p, LpSettings, c_double_p, c_double_p, C.POINTER(LpResult)])
        _declare(lib, "sk_gpu_pdlp_params_default", None, [C.POINTER(PdlpParams)])
        _declare(lib, "sk_gpu_pdlp_solve", C.c_int,
                 [C.c_int, C.c_int, C.c_int, c_int_p, c_int_p, c_double_p,
                  c_int_p, c_int_p, c_double_p, C.c_int,
                  c_double_p, c_double_p, c_double_p, c_double_p, c_double_p,
                  c_double_p, c_double_p, C.POINTER(PdlpParams),
                  c_double_p, c_double_p, C.POINTER(PdlpStats)])
        _declare(lib, "sk_gpu_solver_options_default", None, [C.POINTER(GpuOptions)])
        _declare(lib, "sk_gpu_solve", C.c_int, [M, C.POINTER(SkOptions), C.POINTER(GpuOptions), S,
                                                C.POINTER(PdlpStats)])
        _declare(lib, "sk_gpu_verify", C.c_int, [M, S])
        _declare(lib, "sk_gpu_relative_kkt", C.c_int,
                 [C.c_int, C.c_int, C.c_int, c_int_p, c_int_p, c_double_p, c_double_p,
                  c_double_p, c_double_p, c_double_p, c_double_p, c_double_p, c_double_p,
                  c_double_p])
        vp = C.c_void_p
        _declare(lib, "sk_gpu_pdlp_session_create", C.c_int,
                 [C.c_int, C.c_int, C.c_int, c_int_p, c_int_p, c_double_p, c_double_p, c_double_p,
                  c_double_p, c_double_p, c_double_p, C.POINTER(PdlpParams), C.POINTER(vp)])
        _declare(lib, "sk_gpu_pdlp_session_update", C.c_int,
                 [vp, c_double_p, c_double_p, c_double_p, c_double_p, c_double_p])
        _declare(lib, "sk_gpu_pdlp_session_iterate", C.c_int, [vp, C.c_int, c_double_p])
        _declare(lib, "sk_gpu_pdlp_session_solve", C.c_int,
                 [vp, c_double_p, c_double_p, c_double_p, c_double_p, C.POINTER(PdlpStats)])
        _declare(lib, "sk_gpu_pdlp_session_destroy", None, [vp])
        _declare(lib, "sk_gpu_pdlp_session_set_limits", C.c_int, [vp, C.c_int, C.c_double, C.c_double])
        _declare(lib, "sk_gpu_pdlp_session_resume", C.c_int, [vp, c_double_p, c_double_p, C.POINTER(PdlpStats)])
        _declare(lib, "sk_gpu_pdlp_session_solve_fallback", C.c_int,
                 [vp, c_double_p, c_double_p, C.c_int, c_double_p, c_double_p, C.POINTER(PdlpStats),
                  C.POINTER(C.c_int)])
        _declare(lib, "sk_gpu_gather_rate", C.c_int, [C.c_int, C.c_int, C.c_int, c_double_p])
        _declare(lib, "sk_gpu_read_mps", C.c_int, [C.c_char_p, C.POINTER(SkModel)])
        _declare(lib, "sk_gpu_model_free", None, [C.POINTER(SkModel)])
        _declare(lib, "sk_gpu_batch_options_default", None, [C.POINTER(BatchOptions)])
        _declare(lib, "sk_gpu_batch_solve", C.c_int,
                 [C.c_int, M, C.POINTER(BatchOptions), c_double_p, c_double_p,
                  C.POINTER(BatchResult), C.POINTER(BatchStats)])
        F = C.POINTER(C.c_float)
        _declare(lib, "sk_gpu_mlp_create", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_ulonglong, C.POINTER(vp)])
        _declare(lib, "sk_gpu_mlp_destroy", None, [vp])
        _declare(lib, "sk_gpu_mlp_param_count", C.c_longlong, [vp])
        _declare(lib, "sk_gpu_mlp_set_scaling", C.c_int, [vp, F, F, F, F])
        _declare(lib, "sk_gpu_mlp_get_params", C.c_int, [vp, F])
        _declare(lib, "sk_gpu_mlp_set_params", C.c_int, [vp, F])
        _declare(lib, "sk_gpu_mlp_train", C.c_int, [vp, F, F, C.c_int, C.c_int, C.c_int, C.c_float,
                                                    C.c_ulonglong, F])
        _declare(lib, "sk_gpu_mlp_predict", C.c_int, [vp, F, C.c_int, F])
        _declare(lib, "sk_gpu_mlp_gradients", C.c_int, [vp, F, F, C.c_int, F, F])
        _declare(lib, "sk_gpu_relative_kkt_qp", C.c_int,
                 [C.c_int, C.c_int, C.c_int, c_int_p, c_int_p, c_double_p, c_int_p, c_int_p, c_double_p,
                  C.c_int, c_double_p, c_double_p, c_double_p, c_double_p, c_double_p, c_double_p,
                  c_double_p, c_double_p])
        _declare(lib, "sk_gpu_pdlp_session_create_qp", C.c_int,
                 [C.c_int, C.c_int, C.c_int, c_int_p, c_int_p, c_double_p, c_int_p, c_int_p, c_double_p,
                  C.c_int, c_double_p, c_double_p, c_double_p, c_double_p, c_double_p,
                  C.POINTER(PdlpParams), C.POINTER(vp)])
        _declare(lib, "sk_gpu_model_validate", C.c_int, [M])
        _declare(lib, "sk_gpu_solution_init", None, [S])
        _declare(lib, "sk_gpu_solution_free", None, [S])
        _declare(lib, "sk_gpu_options_default", None, [C.POINTER(SkOptions)])
        _declare(lib, "sk_gpu_result_name", C.c_char_p, [C.c_int])
        LP, PR = C.POINTER(SkLpProblem), C.POINTER(SkPresolveResult)
        _declare(lib, "sk_gpu_presolve", C.c_int, [LP, C.c_int, PR])
        _declare(lib, "sk_gpu_postsolve", C.c_int, [PR, c_double_p, C.c_int, c_double_p])
        _declare(lib, "sk_gpu_presolve_free", None, [PR])
        _declare(lib, "sk_gpu_lp_problem_from_model", C.c_int, [M, LP])
        _declare(lib, "sk_gpu_lp_problem_free", None, [LP])
        _declare(lib, "sk_gpu_lp_solver_options_default", None, [C.POINTER(LpSolverOptions)])
        _declare(lib, "sk_gpu_solve_lp", C.c_int, [LP, C.POINTER(LpSolverOptions), c_double_p,
                                                   C.POINTER(LpSolution)])
        _declare(lib, "sk_gpu_lp_primal_residual", C.c_double, [LP, c_double_p])
        _declare(lib, "sk_gpu_status_name", C.c_char_p, [C.c_int])
        _gpu = lib
    return _gpu


def last_error() -> str:
    return load_gpu().sk_gpu_last_error().decode()


# ---------------------------------------------------------------- models

class Model:
    """An sk_model in CSC form, owned either by numpy arrays or by the
    workload generator (freed with skw_model_free)."""

    def __init__(self):
        self.m = SkModel()
        self.m.objsense = 1.0
        self._keep = []
        self._generated = False
        self._qcsc = None

    @classmethod
    def from_arrays(cls, nrow, ncol, p, i, x, c, rlow, rupp, clow, cupp, q=None):
        """CSC arrays (int32 p[ncol+1], i[nnz], float64 x[nnz]); q is an
        optional (p, i, x) triple for the full symmetric Q."""
        self = cls()
        arrs = [np.ascontiguousarray(p, np.int32), np.ascontiguousarray(i, np.int32),
                np.ascontiguousarray(x, np.float64), np.ascontiguousarray(c, np.float64),
                np.ascontiguousarray(rlow, np.float64), np.ascontiguousarray(rupp, np.float64),
                np.ascontiguousarray(clow, np.float64), np.ascontiguousarray(cupp, np.float64),
                np.zeros(max(ncol, 1), np.uint8)]
        # never hand out zero-length buffers (NULL-looking pointers)
        arrs = [a if a.size else np.zeros(1, a.dtype) for a in arrs]
        self._keep = arrs
        m = self.m
        m.nrow, m.ncol = nrow, ncol
        m.A = SkCsc(nrow, ncol, iptr(arrs[0]), iptr(arrs[1]), dptr(arrs[2]), int(arrs[1].size))
        m.c, m.rlow, m.rupp, m.clow, m.cupp = (dptr(a) for a in arrs[3:8])
        m.vartype = arrs[8].ctypes.data_as(C.POINTER(C.c_ubyte))
        if q is not None:
            qp, qi, qx = (np.ascontiguousarray(q[0], np.int32), np.ascontiguousarray(q[1], np.int32),
                          np.ascontiguousarray(q[2], np.float64))
            qi = qi if qi.size else np.zeros(1, np.int32)
            qx = qx if qx.size else np.zeros(1, np.float64)
            self._keep += [qp, qi, qx]
            self._qcsc = SkCsc(ncol, ncol, iptr(qp), iptr(qi), dptr(qx), int(qi.size))
            m.Q = C.pointer(self._qcsc)
        return self

    @classmethod
    def generated(cls, kind, *args):
        self = cls()
        w = load_workloads()
        rc = getattr(w, kind)(C.byref(self.m), *args)
        if rc != 0:
            raise RuntimeError(f"{kind}{args} failed")
        self._generated = True
        return self

    @classmethod
    def read_mps(cls, path):
        """An MPS / QPS file through the library's reader (sk_gpu_read_mps)."""
        self = cls()
        lib = load_gpu()
        rc = lib.sk_gpu_read_mps(str(path).encode(), C.byref(self.m))
        if rc != 0:
            raise RuntimeError(f"sk_gpu_read_mps({path}) = {rc}: {last_error()}")
        self._owned_by_lib = True
        return self

    def __del__(self):
        if self._generated and _work is not None:
            _work.skw_model_free(C.byref(self.m))
            self._generated = False
        if getattr(self, "_owned_by_lib", False) and _gpu is not None:
            _gpu.sk_gpu_model_free(C.byref(self.m))
            self._owned_by_lib = False

    # numpy views ---------------------------------------------------
    @property
    def nrow(self):
        return self.m.nrow

    @property
    def ncol(self):
        return self.m.ncol

    @property
    def nnz(self):
        return int(self.m.A.p[self.m.ncol])

    def _view(self, ptr, n, dtype=np.float64):
        if n == 0:
            return np.zeros(0, dtype)
        return np.ctypeslib.as_array(ptr, shape=(n,)).view(dtype)

    def csc(self):
        n = self.ncol
        p = self._view(self.m.A.p, n + 1, np.int32).copy()
        nnz = int(p[-1])
        return p, self._view(self.m.A.i, nnz, np.int32).copy(), self._view(self.m.A.x, nnz).copy()

    def q_csc(self):
        if not self.m.Q:
            return None
        q = self.m.Q.contents
        p = self._view(q.p, self.ncol + 1, np.int32).copy()
        nnz = int(p[-1])
        return p, self._view(q.i, nnz, np.int32).copy(), self._view(q.x, nnz).copy()

    def vec(self, name):
        n = self.ncol if name in ("c", "clow", "cupp") else self.nrow
        return self._view(getattr(self.m, name), n).copy()

    def ptr(self):
        return C.byref(self.m)


def csc_to_csr(nrow, ncol, p, i, x):
    """Counting-sort transpose: CSC(A) -> CSR(A) (apps/sankhya_qp_cuda.cu:15-33)."""
    nnz = int(p[-1])
    counts = np.bincount(i[:nnz], minlength=nrow)
    rp = np.zeros(nrow + 1, np.int32)
    np.cumsum(counts, out=rp[1:])
    cols = np.repeat(np.arange(ncol, dtype=np.int32), np.diff(p))
    order = np.argsort(i[:nnz], kind="stable")
    return rp, cols[order].astype(np.int32), x[:nnz][order].astype(np.float64)


# ------------------------------------------------- device memory (tests)

_hip = None


def _hip_runtime():
    """The HIP runtime libsankhya_gpu.so itself links (libamdhip64.so.7), so
    buffers made here live in the same runtime as the library's work."""
    global _hip
    if _hip is None:
        load_gpu()
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    return _hip


class DeviceArray:
    """A float64 device buffer (hipMalloc) with host round trips."""

    H2D, D2H, D2D = 1, 2, 3

    def __init__(self, host):
        host = np.ascontiguousarray(host, np.float64)
        self.n = host.size
        self.ptr = C.c_void_p()
        hip = _hip_runtime()
        if hip.hipMalloc(C.byref(self.ptr), max(8, 8 * self.n)) != 0:
            raise RuntimeError("hipMalloc failed")
        if self.n and hip.hipMemcpy(self.ptr, host.ctypes.data, 8 * self.n, self.H2D) != 0:
            raise RuntimeError("hipMemcpy H2D failed")

    @property
    def value(self):
        return self.ptr.value

    def numpy(self):
        out = np.empty(self.n)
        hip = _hip_runtime()
        hip.hipDeviceSynchronize()
        if self.n and hip.hipMemcpy(out.ctypes.data, self.ptr, 8 * self.n, self.D2H) != 0:
            raise RuntimeError("hipMemcpy D2H failed")
        return out

    def __del__(self):
        if self.ptr and _hip is not None:
            _hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def device_synchronize():
    _hip_runtime().hipDeviceSynchronize()


def batch_solve(models, block_threads=0, use_graph=1, **pdlp):
    """sk_gpu_batch_solve over a list of Model objects; returns
    (rc, [x_k], [y_k], [BatchResult], BatchStats)."""
    lib = load_gpu()
    opt = BatchOptions()
    lib.sk_gpu_batch_options_default(C.byref(opt))
    opt.block_threads = block_threads
    opt.use_graph = use_graph
    for k, v in pdlp.items():
        setattr(opt.pdlp, k, v)
    arr = (SkModel * len(models))()
    for k, md in enumerate(models):
        arr[k] = md.m
    cols = sum(md.ncol for md in models)
    rows = sum(md.nrow for md in models)
    x = np.zeros(max(cols, 1))
    y = np.zeros(max(rows, 1))
    res = (BatchResult * max(len(models), 1))()
    st = BatchStats()
    rc = lib.sk_gpu_batch_solve(len(models), arr, C.byref(opt), dptr(x), dptr(y), res, C.byref(st))
    xs, ys, c0, r0 = [], [], 0, 0
    for md in models:
        xs.append(x[c0:c0 + md.ncol].copy()); c0 += md.ncol
        ys.append(y[r0:r0 + md.nrow].copy()); r0 += md.nrow
    return rc, xs, ys, [res[k] for k in range(len(models))], st


def fptr(a):
    return _ptr(a, C.c_float) if a is not None else None


class Mlp:
    """The warm-start MLP on the GPU (sk_gpu_mlp_*)."""

    def __init__(self, n_in, hidden, n_out, seed=1):
        self.lib = load_gpu()
        self.h = C.c_void_p()
        self.shape = (n_in, hidden, n_out)
        if self.lib.sk_gpu_mlp_create(n_in, hidden, n_out, seed, C.byref(self.h)) != 0:
            raise RuntimeError(last_error())
        self.n_params = int(self.lib.sk_gpu_mlp_param_count(self.h))

    def _f32(self, a):
        return np.ascontiguousarray(a, np.float32)

    def set_scaling(self, in_mean, in_std, out_mean, out_std):
        arrs = [self._f32(a) for a in (in_mean, in_std, out_mean, out_std)]
        if self.lib.sk_gpu_mlp_set_scaling(self.h, *[fptr(a) for a in arrs]) != 0:
            raise RuntimeError(last_error())

    def params(self):
        out = np.zeros(self.n_params, np.float32)
        assert self.lib.sk_gpu_mlp_get_params(self.h, fptr(out)) == 0
        return out

    def set_params(self, p):
        p = self._f32(p)
        assert self.lib.sk_gpu_mlp_set_params(self.h, fptr(p)) == 0

    def gradients(self, X, Y):
        X, Y = self._f32(X), self._f32(Y)
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_float()
        rc = self.lib.sk_gpu_mlp_gradients(self.h, fptr(X), fptr(Y), X.shape[0], C.byref(loss), fptr(g))
        if rc != 0:
            raise RuntimeError(last_error())
        return float(loss.value), g

    def train(self, X, Y, epochs, batch, lr, seed=7):
        X, Y = self._f32(X), self._f32(Y)
        hist = np.zeros(max(epochs, 1), np.float32)
        rc = self.lib.sk_gpu_mlp_train(self.h, fptr(X), fptr(Y), X.shape[0], epochs, batch, lr, seed, fptr(hist))
        if rc != 0:
            raise RuntimeError(last_error())
        return hist[:epochs]

    def predict(self, X):
        X = self._f32(X)
        Y = np.zeros((X.shape[0], self.shape[2]), np.float32)
        if self.lib.sk_gpu_mlp_predict(self.h, fptr(X), X.shape[0], fptr(Y)) != 0:
            raise RuntimeError(last_error())
        return Y

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.sk_gpu_mlp_destroy(self.h)
            self.h = C.c_void_p()


# ---------------------------------------------------------------- LP pipeline

class SkLpProblem(C.Structure):
    """sk_gpu_lp_problem: the reference's LpProblem (src/lp/LpProblem.hpp:15-28)."""
    _fields_ = [("rows", C.c_int), ("cols", C.c_int), ("nnz", C.c_int),
                ("row_ptr", c_int_p), ("col_idx", c_int_p), ("values", c_double_p),
                ("obj", c_double_p), ("rhs", c_double_p), ("row_types", C.POINTER(C.c_char)),
                ("lower", c_double_p), ("upper", c_double_p),
                ("slack_lower", c_double_p), ("slack_upper", c_double_p)]


class SkPresolveResult(C.Structure):
    """sk_gpu_presolve_result: PresolveResult (src/lp/Presolve.hpp:12-55)."""
    _fields_ = [("status", C.c_int), ("reason", C.c_char_p), ("reason_row", C.c_int),
                ("reason_col", C.c_int), ("reason_pass", C.c_int), ("reduced", SkLpProblem),
                ("kept_columns", c_int_p), ("kept_rows", c_int_p), ("fixed_value", c_double_p),
                ("column_removed", C.POINTER(C.c_char)), ("original_n_rows", C.c_int),
                ("original_n_cols", C.c_int), ("removed_rows", C.c_int), ("removed_cols", C.c_int)]


class LpSolverOptions(C.Structure):
    _fields_ = [("use_presolve", C.c_int), ("presolve_passes", C.c_int), ("pdlp", PdlpParams)]


class LpSolution(C.Structure):
    _fields_ = [("status", C.c_int), ("has_x", C.c_int), ("objective_value", C.c_double),
                ("iterations", C.c_int), ("presolve_seconds", C.c_double),
                ("solve_seconds", C.c_double), ("presolve_removed_rows", C.c_int),
                ("presolve_removed_cols", C.c_int), ("used_first_order", C.c_int),
                ("primal_residual", C.c_double), ("pdlp", PdlpStats)]


PRESOLVE_OK, PRESOLVE_INFEASIBLE, PRESOLVE_UNBOUNDED = 0, 1, 2
LP_OPTIMAL, LP_INFEASIBLE, LP_UNBOUNDED, LP_ITERATION_LIMIT, LP_NUMERICAL_FAILURE = 0, 1, 2, 3, 4
LP_STATUS_NAMES = ["optimal", "infeasible", "unbounded", "iteration_limit", "numerical_failure"]
_LP_FIELDS = ("row_ptr", "col_idx", "values", "obj", "rhs", "row_types", "lower", "upper",
              "slack_lower", "slack_upper")
_LP_DTYPES = {"row_ptr": np.int32, "col_idx": np.int32, "row_types": np.uint8}


class LpProblem:
    """An LpProblem held in numpy arrays: A x + s = rhs, slack_lower <= s <=
    slack_upper, lower <= x <= upper, A in CSR, absent bounds IEEE inf."""

    def __init__(self, rows, cols, row_ptr, col_idx, values, obj, rhs, lower, upper,
                 slack_lower, slack_upper, row_types=None):
        self.rows, self.cols = int(rows), int(cols)
        a = {"row_ptr": row_ptr, "col_idx": col_idx, "values": values, "obj": obj, "rhs": rhs,
             "lower": lower, "upper": upper, "slack_lower": slack_lower, "slack_upper": slack_upper,
             "row_types": row_types if row_types is not None else np.full(self.rows, ord("L"), np.uint8)}
        for k, v in a.items():
            if k == "row_types" and isinstance(v, (bytes, str)):
                v = np.frombuffer(v.encode() if isinstance(v, str) else v, np.uint8)
            setattr(self, k, np.ascontiguousarray(v, _LP_DTYPES.get(k, np.float64)))
        self.nnz = int(self.row_ptr[-1]) if len(self.row_ptr) else 0

    @classmethod
    def from_struct(cls, s):
        """Copy an SkLpProblem (library- or reference-owned) into numpy arrays."""
        m, n, nnz = s.rows, s.cols, s.nnz
        def arr(ptr, k, dt):
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(dt)), shape=(k,)).copy() if k and ptr else np.zeros(0, dt)
        return cls(m, n, arr(s.row_ptr, m + 1, C.c_int) if s.row_ptr else np.zeros(m + 1, np.int32),
                   arr(s.col_idx, nnz, C.c_int), arr(s.values, nnz, C.c_double), arr(s.obj, n, C.c_double),
                   arr(s.rhs, m, C.c_double), arr(s.lower, n, C.c_double), arr(s.upper, n, C.c_double),
                   arr(s.slack_lower, m, C.c_double), arr(s.slack_upper, m, C.c_double),
                   arr(s.row_types, m, C.c_uint8) if s.row_types else None)

    @classmethod
    def from_model(cls, model):
        """sk_gpu_lp_problem_from_model of a Model (continuous LP)."""
        lib = load_gpu()
        s = SkLpProblem()
        if lib.sk_gpu_lp_problem_from_model(model.ptr(), C.byref(s)) != 0:
            raise RuntimeError(last_error())
        out = cls.from_struct(s)
        lib.sk_gpu_lp_problem_free(C.byref(s))
        return out

    def struct(self):
        s = SkLpProblem(self.rows, self.cols, self.nnz)
        for k in _LP_FIELDS:
            v = getattr(self, k)
            ct = {"row_ptr": C.c_int, "col_idx": C.c_int, "row_types": C.c_char}.get(k, C.c_double)
            setattr(s, k, v.ctypes.data_as(C.POINTER(ct)) if v.size else None)
        return s

    def arrays(self):
        return {k: getattr(self, k) for k in _LP_FIELDS} | {"rows": self.rows, "cols": self.cols}


def presolve_result_dict(r, free=None):
    """Every field of an SkPresolveResult as numpy arrays (then freed with
    `free`, the owning library's release function)."""
    n = r.original_n_cols
    ok = r.status == PRESOLVE_OK
    kc = r.original_n_cols - r.removed_cols if ok else 0
    kr = r.original_n_rows - r.removed_rows if ok else 0
    def arr(ptr, k, dt):
        return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(dt)), shape=(k,)).copy() if k and ptr else np.zeros(0, dt)
    d = {"status": r.status, "reason": r.reason.decode() if r.reason else None,
         "reason_row": r.reason_row, "reason_col": r.reason_col, "reason_pass": r.reason_pass,
         "original_n_rows": r.original_n_rows, "original_n_cols": n,
         "removed_rows": r.removed_rows, "removed_cols": r.removed_cols,
         "kept_columns": arr(r.kept_columns, kc, C.c_int), "kept_rows": arr(r.kept_rows, kr, C.c_int),
         "fixed_value": arr(r.fixed_value, n, C.c_double),
         "column_removed": arr(r.column_removed, n, C.c_uint8),
         "reduced": LpProblem.from_struct(r.reduced) if ok else None}
    if free is not None:
        free(C.byref(r))
    return d


def presolve(problem: LpProblem, max_passes=20):
    """sk_gpu_presolve (src/lp/Presolve.hpp:92) as a dict of arrays."""
    lib = load_gpu()
    s = problem.struct()
    r = SkPresolveResult()
    if lib.sk_gpu_presolve(C.byref(s), max_passes, C.byref(r)) != 0:
        raise RuntimeError(last_error())
    return presolve_result_dict(r, lib.sk_gpu_presolve_free)


def postsolve(problem: LpProblem, reduced_x, max_passes=20):
    """sk_gpu_postsolve of sk_gpu_presolve(problem) applied to reduced_x."""
    lib = load_gpu()
    s = problem.struct()
    r = SkPresolveResult()
    if lib.sk_gpu_presolve(C.byref(s), max_passes, C.byref(r)) != 0:
        raise RuntimeError(last_error())
    rx = np.ascontiguousarray(reduced_x, np.float64)
    x = np.zeros(max(problem.cols, 1))
    rc = lib.sk_gpu_postsolve(C.byref(r), dptr(rx) if rx.size else None, rx.size, dptr(x))
    lib.sk_gpu_presolve_free(C.byref(r))
    if rc != 0:
        raise RuntimeError(last_error())
    return x[:problem.cols]


def lp_solver_options(use_presolve=1, presolve_passes=20, **pdlp) -> LpSolverOptions:
    return LpSolverOptions(use_presolve=use_presolve, presolve_passes=presolve_passes,
                           pdlp=default_pdlp_params(**pdlp))


def solve_lp(problem: LpProblem, options: LpSolverOptions | None = None):
    """sk_gpu_solve_lp (src/lp/LpSolver.cpp:127, FIRST_ORDER): (solution, x)."""
    lib = load_gpu()
    s = problem.struct()
    o = options if options is not None else lp_solver_options()
    x = np.zeros(max(problem.cols, 1))
    sol = LpSolution()
    if lib.sk_gpu_solve_lp(C.byref(s), C.byref(o), dptr(x), C.byref(sol)) != 0:
        raise RuntimeError(last_error())
    return sol, x[:problem.cols]


def primal_residual(problem: LpProblem, x):
    lib = load_gpu()
    s = problem.struct()
    xs = np.ascontiguousarray(x, np.float64)
    return lib.sk_gpu_lp_primal_residual(C.byref(s), dptr(xs) if xs.size else None)
