"""ctypes bindings of the MI355X presolve engine (libgpupre.so) and of the
synthetic LP generator (liblpgen.so).

The engine's C ABI is declared in include/gpupre.h.  This module is plumbing
for tests, the bench and Python callers: it mirrors the C structs, uploads
host LPs and exposes preprocess / postprocess / free_summary with the
reference's semantics (gpu_preprocessor.hpp:22-39).  It never falls back to
CPU code: if the shared library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgpupre.so")
LPGEN_PATH = os.path.join(HERE, "liblpgen.so")

OFFSET = C.c_int32  # gp_offset (NnzOffset, int32 build)
NP_OFFSET = np.int32

RULES = [
    "fixed_variables", "empty_columns", "empty_rows", "singleton_rows", "singleton_columns",
    "bound_strengthening", "redundant_rows", "dual_fix", "doubleton_equations",
    "parallel_rows", "parallel_columns",
]
RULE = {name: i for i, name in enumerate(RULES)}
# structural Startup rules (GP_STARTUP_*), in the order Startup runs them
STARTUP_RULES = ["column_domination", "two_row_elimination", "zero_cost_products"]
STARTUP_RULE = {name: i for i, name in enumerate(STARTUP_RULES)}
# GP_PRIMAL_ONLY_* bits of gp_summary.primal_only_reductions
PRIMAL_ONLY_COLUMN_DOMINATION = 1
PRIMAL_ONLY_PROJECTED_AUXILIARY = 2
PRIMAL_ONLY_VARIABLE_PAIR = 4
# postsolve record types (the GP_RED_* enum of gpupre.h)
RED_FIXED_VARIABLE = 0
RED_INFINITE_BOUND_DUAL_FIX = 1
RED_SINGLETON_COLUMN_SUBSTITUTION = 2
RED_PARALLEL_COLUMN = 3
RED_PARALLEL_ROW = 4
RED_DELETED_ROW = 5
RED_ROW_LOWER_BOUND_CHANGE = 8
RED_ROW_UPPER_BOUND_CHANGE = 9
RED_ROW_DUAL_SHIFT = 10
RED_VARIABLE_BOUND_CHANGE = 11
RED_ROW_IMPLIED_BOUND_CHANGE = 12
RED_DOUBLETON_EQUATION = 13
RED_FME_COL = 14
# structural primal recovery kinds (GP_RECOVERY_* of gpupre.h)
RECOVERY_ABS_VALUE = 0
RECOVERY_VAR_PAIR = 1
PHASES = ["cleanup", "propagation", "aggregation", "parallel", "end"]
KERNEL_FAMILIES = ["live_stats", "activity", "bound_candidates", "compact", "transpose",
                   "rules_other", "hash_sort", "sc_activity", "redundant", "col_max", "live_stats_inc",
                   "cleanup", "singleton_columns", "bounds", "dual_fix", "doubleton", "parallel", "vectors",
                   "entry_bounds"]


class GpCsr(C.Structure):
    _fields_ = [("rows", C.c_int32), ("cols", C.c_int32), ("nnz", OFFSET),
                ("row_ptr", C.c_void_p), ("col_idx", C.c_void_p), ("vals", C.c_void_p)]


class GpDeviceLp(C.Structure):
    _fields_ = [("A", GpCsr), ("AT", GpCsr), ("c", C.c_void_p), ("AL", C.c_void_p),
                ("AU", C.c_void_p), ("l", C.c_void_p), ("u", C.c_void_p),
                ("obj_constant", C.c_double)]


class GpMpsModel(C.Structure):
    _fields_ = [("rows", C.c_int32), ("cols", C.c_int32), ("entries", C.c_int64), ("entry_row", C.c_void_p),
                ("entry_col", C.c_void_p), ("entry_val", C.c_void_p), ("c", C.c_void_p), ("AL", C.c_void_p),
                ("AU", C.c_void_p), ("l", C.c_void_p), ("u", C.c_void_p), ("obj_constant", C.c_double),
                ("name", C.c_void_p)]


class GpLpSolveParams(C.Structure):
    _fields_ = [("struct_size", C.c_int32), ("max_iterations", C.c_int32), ("tolerance", C.c_double),
                ("time_limit", C.c_double), ("check_every", C.c_int32), ("ruiz_iterations", C.c_int32),
                ("pock_chambolle", C.c_int32), ("bound_cost_scaling", C.c_int32), ("power_iterations", C.c_int32),
                ("reserved", C.c_int32), ("power_tolerance", C.c_double), ("restart_sufficient", C.c_double),
                ("restart_necessary", C.c_double), ("restart_artificial", C.c_double),
                ("residual_balance", C.c_double), ("residual_balance_band", C.c_double)]


class GpLpSolveInfo(C.Structure):
    _fields_ = [("status", C.c_int32), ("iterations", C.c_int32), ("restarts", C.c_int32), ("reserved", C.c_int32),
                ("primal_objective", C.c_double), ("dual_objective", C.c_double),
                ("primal_residual", C.c_double), ("dual_residual", C.c_double), ("gap", C.c_double),
                ("seconds", C.c_double), ("sigma", C.c_double), ("lambda_", C.c_double),
                ("bytes_per_iteration", C.c_double)]


LP_STATUS = {0: "OPTIMAL", 1: "ITER_LIMIT", 2: "TIME_LIMIT", 3: "NUMERICAL_ERROR", 4: "UNBOUNDED", 5: "INFEASIBLE"}


class GpFoldResult(C.Structure):
    _fields_ = [("applied", C.c_int32), ("original_rows", C.c_int32), ("original_cols", C.c_int32),
                ("folded_rows", C.c_int32), ("folded_cols", C.c_int32), ("rounds", C.c_int32),
                ("refinement_ok", C.c_int32), ("final_row_colors", C.c_int32), ("final_col_colors", C.c_int32),
                ("reserved", C.c_int32), ("seconds", C.c_double), ("folded_lp", GpDeviceLp),
                ("row_color_id", C.c_void_p), ("col_color_id", C.c_void_p), ("row_scale", C.c_void_p),
                ("col_scale", C.c_void_p), ("state", C.c_void_p)]


class GpParams(C.Structure):
    _fields_ = [
        ("struct_size", C.c_int32), ("max_iters", C.c_int32), ("max_time", C.c_double),
        ("verbose", C.c_int32), ("record_postsolve_tape", C.c_int32),
        ("record_postsolve_tape_cpu", C.c_int32), ("enable_presolve", C.c_int32),
        ("enable_folding", C.c_int32), ("folding_tolerance", C.c_double),
        ("feasibility_tol", C.c_double), ("bound_tol", C.c_double), ("zero_tol", C.c_double),
        ("bound_strengthening_upper_relaxation_factor", C.c_double),
        ("bound_strengthening_max_rounds", C.c_int32),
        ("bound_strengthening_max_bound_only_rounds", C.c_int32),
        ("bound_strengthening_preserve_free_zero_cost", C.c_int32),
        ("doubleton_equations_max_fill_in_proxy", C.c_int32),
        ("doubleton_equations_scan", C.c_int32),
        ("doubleton_equations_min_selected_per_batch", C.c_int32),
        ("doubleton_equations_min_selected_ratio", C.c_double),
        ("doubleton_equations_max_batch_rounds", C.c_int32),
        ("doubleton_equations_max_time", C.c_double),
        ("enable_infeasible_and_fixed_variables", C.c_int32), ("enable_empty_rows", C.c_int32),
        ("enable_singleton_rows", C.c_int32), ("enable_infeasible_and_redundant_rows", C.c_int32),
        ("enable_bound_strengthening", C.c_int32), ("enable_parallel_rows", C.c_int32),
        ("enable_empty_columns", C.c_int32), ("enable_singleton_columns", C.c_int32),
        ("enable_doubleton_equations", C.c_int32), ("enable_dual_fix", C.c_int32),
        ("dual_fix_max_nonzero_cost_col_degree", C.c_int32),
        ("enable_parallel_columns", C.c_int32), ("parallel_columns_max_abs_ratio", C.c_double),
        ("parallel_columns_scale_guard_max_group_size", C.c_int32),
        ("parallel_columns_scale_guard_min_groups", C.c_int32),
        ("enable_structure_specific_rules", C.c_int32),
        ("enable_two_row_variable_elimination", C.c_int32),
        ("two_row_variable_elimination_max_row_nnz", C.c_int32),
        ("two_row_variable_elimination_safety_tol", C.c_double),
        ("two_row_variable_elimination_min_candidates", C.c_int32),
        ("two_row_variable_elimination_min_candidate_ratio", C.c_double),
        ("enable_zero_cost_product_variables", C.c_int32),
        ("zero_cost_product_variables_min_candidates", C.c_int32),
        ("zero_cost_product_variables_min_candidate_ratio", C.c_double),
        ("enable_piecewise_linear_reformulation_variable_pair", C.c_int32),
        ("piecewise_linear_reformulation_variable_pair_epigraph_probe_pairs", C.c_int32),
        ("piecewise_linear_reformulation_variable_pair_epigraph_min_pairs", C.c_int32),
        ("piecewise_linear_reformulation_variable_pair_epigraph_min_potential_gain", C.c_double),
        ("piecewise_linear_reformulation_variable_pair_epigraph_min_actual_gain", C.c_double),
        ("piecewise_linear_reformulation_variable_pair_epigraph_max_tail_cols", C.c_int32),
        ("piecewise_linear_reformulation_variable_pair_epigraph_max_quotient_terms_per_row", C.c_int32),
        ("piecewise_linear_reformulation_variable_pair_shifted_antipodal_probe_pairs", C.c_int32),
        ("piecewise_linear_reformulation_variable_pair_shifted_antipodal_min_pairs", C.c_int32),
        ("piecewise_linear_reformulation_variable_pair_shifted_antipodal_min_potential_gain", C.c_double),
        ("piecewise_linear_reformulation_variable_pair_shifted_antipodal_min_actual_gain", C.c_double),
        ("piecewise_linear_reformulation_variable_pair_shifted_antipodal_max_quotient_terms_per_row", C.c_int32),
        ("enable_column_domination", C.c_int32),
        ("column_domination_probe_rows", C.c_int32),
        ("column_domination_probe_cols", C.c_int32),
        ("column_domination_probe_nnz", C.c_int32),
        ("column_domination_min_fixed_cols", C.c_int32),
        ("column_domination_min_candidate_ratio", C.c_double),
        ("enable_piecewise_linear_reformulation_absolute_value", C.c_int32),
        ("piecewise_linear_reformulation_absolute_value_residual_bound_mode", C.c_int32),
        ("piecewise_linear_reformulation_absolute_value_residual_bound_as_free_min", C.c_double),
        ("enable_redundant_variable_bounds", C.c_int32),
        ("enable_startup", C.c_int32),
        ("bound_strengthening_bound_only_nnz_round_budget", C.c_int64),
        ("collect_trace", C.c_int32), ("learned_schedule", C.c_int32),
        ("scheduler_weights", C.c_void_p), ("kernel_timing", C.c_int32),
        ("incremental_live_stats", C.c_int32),
    ]


class GpRoundTrace(C.Structure):
    _fields_ = [("phase", C.c_int32), ("next_phase", C.c_int32), ("nnz_before", C.c_int64),
                ("nnz_after", C.c_int64), ("rows_live", C.c_int32), ("cols_live", C.c_int32),
                ("fired_mask", C.c_int32), ("skipped_mask", C.c_int32), ("seconds", C.c_double),
                ("bound_only_rounds", C.c_int32), ("bound_only_limit", C.c_int32), ("final_sweep", C.c_int32),
                ("reserved", C.c_int32)]


GP_KT_COUNT = 19

# learned pass scheduler network (gpupre.h GP_MLP_*)
GP_MLP_IN, GP_MLP_H1, GP_MLP_H2, GP_MLP_OUT = 32, 64, 64, 7
GP_MLP_PARAMS = GP_MLP_H1 * GP_MLP_IN + GP_MLP_H1 + GP_MLP_H2 * GP_MLP_H1 + GP_MLP_H2 + GP_MLP_OUT * GP_MLP_H2 + GP_MLP_OUT
GP_MLP_BLOB = GP_MLP_PARAMS + GP_MLP_OUT


class GpSummary(C.Structure):
    _fields_ = [
        ("preprocessing_seconds", C.c_double), ("workspace_setup_seconds", C.c_double),
        ("workspace_release_seconds", C.c_double), ("presolve_enabled", C.c_int32),
        ("folding_applied", C.c_int32), ("original_rows", C.c_int32), ("original_cols", C.c_int32),
        ("reduced_rows", C.c_int32), ("reduced_cols", C.c_int32), ("reduced_nnz", OFFSET),
        ("iterations", C.c_int32), ("stopped_at_iteration_limit", C.c_int32),
        ("has_change", C.c_int32), ("elapsed_seconds", C.c_double),
        ("obj_constant_delta", C.c_double), ("has_infeasible", C.c_int32),
        ("has_unbounded", C.c_int32), ("reduced_lp", GpDeviceLp), ("owns_reduced_lp", C.c_int32),
        ("record", C.c_void_p), ("trace_count", C.c_int32), ("trace", C.POINTER(GpRoundTrace)),
        ("kernel_ms", C.c_double * GP_KT_COUNT), ("kernel_launches", C.c_int64 * GP_KT_COUNT),
        ("kernel_bytes", C.c_double * GP_KT_COUNT), ("unsupported_stages", C.c_int32),
        ("folding", GpFoldResult), ("primal_only_reductions", C.c_int32),
    ]


class GpBatchSummary(C.Structure):
    _fields_ = [("count", C.c_int32), ("whole", GpSummary), ("assembled", GpDeviceLp),
                ("row_offsets", C.POINTER(C.c_int32)), ("col_offsets", C.POINTER(C.c_int32)),
                ("reduced_row_offsets", C.POINTER(C.c_int32)), ("reduced_col_offsets", C.POINTER(C.c_int32)),
                ("reduced_nnz_offsets", C.POINTER(OFFSET)), ("obj_constant", C.POINTER(C.c_double)),
                ("assemble_seconds", C.c_double), ("presolve_seconds", C.c_double), ("total_seconds", C.c_double),
                ("status", C.POINTER(C.c_int32)), ("restarts", C.c_int32)]


class GpTuning(C.Structure):
    _fields_ = [("lanes_candidates", C.c_int32), ("lanes_col_max", C.c_int32), ("compact_per_lane", C.c_int32),
                ("lanes_live_stats", C.c_int32), ("fold_rows", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class GpRecordSizes(C.Structure):
    _fields_ = [("m0", C.c_int32), ("n0", C.c_int32), ("m1", C.c_int32), ("n1", C.c_int32),
                ("fixed_count", C.c_int32), ("record_count", C.c_int64),
                ("index_count", C.c_int64), ("value_count", C.c_int64)]


class GpPlan(C.Structure):
    _fields_ = [
        ("keep_row_mask", C.c_void_p), ("keep_col_mask", C.c_void_p), ("new_c", C.c_void_p),
        ("new_l", C.c_void_p), ("new_u", C.c_void_p), ("new_AL", C.c_void_p),
        ("new_AU", C.c_void_p), ("rule_entry_keep_col", C.c_void_p),
        ("propagation_reference_l", C.c_void_p), ("propagation_reference_u", C.c_void_p),
        ("obj_constant_delta", C.c_double), ("has_change", C.c_int32),
        ("has_row_action", C.c_int32), ("has_col_action", C.c_int32),
        ("has_redundant_row_deletion", C.c_int32), ("has_row_deletion", C.c_int32),
        ("has_infeasible", C.c_int32), ("has_unbounded", C.c_int32), ("has_new_A", C.c_int32),
        ("new_A", GpCsr), ("tape", C.c_void_p), ("rule_cache", C.c_void_p),
    ]


class GpStats(C.Structure):
    _fields_ = [
        ("live_masks", C.c_int32), ("live_nnz", C.c_int64), ("row_nnz", C.c_void_p),
        ("empty_row_mask", C.c_void_p), ("singleton_row_mask", C.c_void_p),
        ("singleton_row_col", C.c_void_p), ("singleton_row_val", C.c_void_p),
        ("col_nnz", C.c_void_p), ("empty_col_mask", C.c_void_p),
        ("singleton_column_mask", C.c_void_p), ("singleton_column_row", C.c_void_p),
        ("singleton_column_val", C.c_void_p),
    ]


PLAN_FLAGS = ["has_change", "has_row_action", "has_col_action", "has_redundant_row_deletion",
              "has_row_deletion", "has_infeasible", "has_unbounded", "has_new_A"]


# ---------------------------------------------------------------- host LP container
@dataclass
class HostLP:
    """CSR LP on the host: min c'x + obj_constant, AL <= Ax <= AU, l <= x <= u."""
    m: int
    n: int
    row_ptr: np.ndarray
    col_idx: np.ndarray
    vals: np.ndarray
    c: np.ndarray
    l: np.ndarray
    u: np.ndarray
    AL: np.ndarray
    AU: np.ndarray
    obj_constant: float = 0.0
    x0: np.ndarray | None = None
    y0: np.ndarray | None = None
    at: tuple | None = field(default=None, repr=False)

    @property
    def nnz(self) -> int:
        return int(self.row_ptr[-1])

    def transpose(self):
        """(row_ptr, col_idx, vals) of A^T with rows of each column in increasing order."""
        if self.at is None:
            rows = np.repeat(np.arange(self.m, dtype=np.int32), np.diff(self.row_ptr))
            order = np.argsort(self.col_idx, kind="stable")
            counts = np.bincount(self.col_idx, minlength=self.n)
            rp = np.zeros(self.n + 1, dtype=NP_OFFSET)
            np.cumsum(counts, out=rp[1:])
            self.at = (rp, rows[order].astype(np.int32), self.vals[order].copy())
        return self.at

    @staticmethod
    def from_dense(A, c, l, u, AL, AU, obj_constant=0.0) -> "HostLP":
        A = np.asarray(A, dtype=np.float64)
        m, n = A.shape
        rp = [0]
        ci, va = [], []
        for i in range(m):
            nz = np.nonzero(A[i])[0]
            ci.extend(nz.tolist())
            va.extend(A[i, nz].tolist())
            rp.append(len(ci))
        f = lambda v, k: np.asarray(v, dtype=np.float64).reshape(k).copy()
        return HostLP(m, n, np.asarray(rp, dtype=NP_OFFSET), np.asarray(ci, dtype=np.int32),
                      np.asarray(va, dtype=np.float64), f(c, n), f(l, n), f(u, n), f(AL, m), f(AU, m),
                      float(obj_constant))


# ---------------------------------------------------------------- library loading
_lib = None
_lpgen = None


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "gp_last_error": (C.c_char_p, []),
        "gp_abi_version": (C.c_int, []),
        "gp_default_params": (None, [C.POINTER(GpParams)]),
        "gp_preprocess": (C.c_int, [C.POINTER(GpDeviceLp), C.POINTER(GpParams), C.POINTER(GpSummary)]),
        "gp_postprocess": (C.c_int, [vp, vp, vp, C.POINTER(GpSummary), C.POINTER(GpDeviceLp), vp, vp, vp]),
        "gp_postsolve_record": (C.c_int, [vp, vp, vp, C.POINTER(GpHostRecord), C.POINTER(GpDeviceLp), vp, vp, vp, vp]),
        "gp_free_summary": (None, [C.POINTER(GpSummary)]),
        "gp_record_get_sizes": (C.c_int, [C.POINTER(GpSummary), C.POINTER(GpRecordSizes)]),
        "gp_record_copy": (C.c_int, [C.POINTER(GpSummary)] + [vp] * 10),
        "gp_plan_create": (C.c_int, [C.POINTER(GpDeviceLp), C.POINTER(GpPlan)]),
        "gp_plan_destroy": (None, [C.POINTER(GpPlan)]),
        "gp_plan_enable_tape": (C.c_int, [C.POINTER(GpPlan)]),
        "gp_stats_create": (C.c_int, [C.c_int32, C.c_int32, C.POINTER(GpStats)]),
        "gp_stats_destroy": (None, [C.POINTER(GpStats)]),
        "gp_refresh_live_stats": (C.c_int, [C.POINTER(GpStats), C.POINTER(GpDeviceLp), C.POINTER(GpPlan)]),
        "gp_update_live_stats": (C.c_int, [C.POINTER(GpStats), C.POINTER(GpDeviceLp), C.POINTER(GpPlan), vp, vp,
                                           C.POINTER(C.c_int32)]),
        "gp_rule_apply": (C.c_int, [C.c_int32, C.POINTER(GpPlan), C.POINTER(GpDeviceLp), C.POINTER(GpStats),
                                     C.POINTER(GpParams)]),
        "gp_startup_rule": (C.c_int, [C.c_int32, C.POINTER(GpDeviceLp), C.POINTER(GpParams), C.POINTER(GpPlan),
                                       C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
        "gp_plan_tape_sizes": (C.c_int, [C.POINTER(GpPlan), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                          C.POINTER(C.c_int64)]),
        "gp_plan_tape_copy": (C.c_int, [C.POINTER(GpPlan)] + [vp] * 6),
        "gp_fixed_variable_tape": (C.c_int, [C.POINTER(GpPlan), vp, vp, vp, vp, vp, C.POINTER(GpCsr), C.c_int32,
                                             C.c_int32, C.POINTER(C.c_double)]),
        "gp_row_activity_summary": (C.c_int, [C.POINTER(GpCsr), vp, vp, C.c_double, vp, vp, vp, vp, vp, vp]),
        "gp_row_activity_summary_noalloc": (C.c_int, [C.POINTER(GpCsr), vp, vp, C.c_double, vp, vp, vp, vp, vp,
                                                      vp]),
        "gp_transpose_csr": (C.c_int, [C.POINTER(GpCsr), C.POINTER(GpCsr)]),
        "gp_parallel_row_hashes": (C.c_int, [C.POINTER(GpCsr), vp, vp, C.c_double, vp]),
        "gp_radix_sort_pairs": (C.c_int, [vp, vp, C.c_int64, C.c_int32]),
        "gp_compact_csr": (C.c_int, [C.POINTER(GpCsr), vp, vp, C.POINTER(GpCsr)]),
        "gp_csr_free": (None, [C.POINTER(GpCsr)]),
        "gp_device_count": (C.c_int, [C.POINTER(C.c_int32)]),
        "gp_malloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t]),
        "gp_free": (C.c_int, [vp]),
        "gp_memcpy_h2d": (C.c_int, [vp, vp, C.c_size_t]),
        "gp_memcpy_d2h": (C.c_int, [vp, vp, C.c_size_t]),
        "gp_memcpy_d2d": (C.c_int, [vp, vp, C.c_size_t]),
        "gp_synchronize": (C.c_int, []),
        "gp_lp_upload": (C.c_int, [C.c_int32, C.c_int32, OFFSET, vp, vp, vp, vp, vp, vp, vp, vp, C.c_double,
                                    C.POINTER(GpDeviceLp)]),
        "gp_lp_free": (None, [C.POINTER(GpDeviceLp)]),
        "gp_lp_upload_entries": (C.c_int, [C.c_int32, C.c_int32, C.c_int64, vp, vp, vp, vp, vp, vp, vp, vp,
                                           C.c_double, C.POINTER(GpDeviceLp)]),
        "gp_mps_parse": (C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(GpMpsModel)]),
        "gp_mps_read": (C.c_int, [C.c_char_p, C.POINTER(GpMpsModel)]),
        "gp_mps_free": (None, [C.POINTER(GpMpsModel)]),
        "gp_mps_upload": (C.c_int, [C.POINTER(GpMpsModel), C.POINTER(GpDeviceLp)]),
        "gp_lp_solve_default_params": (None, [C.POINTER(GpLpSolveParams)]),
        "gp_fold": (C.c_int, [C.POINTER(GpDeviceLp), C.c_double, C.POINTER(GpFoldResult)]),
        "gp_unfold": (C.c_int, [C.POINTER(GpFoldResult), vp, vp, vp, vp, vp, vp]),
        "gp_free_fold": (None, [C.POINTER(GpFoldResult)]),
        "gp_lp_solve": (C.c_int, [C.POINTER(GpDeviceLp), C.POINTER(GpLpSolveParams), vp, vp, vp,
                                  C.POINTER(GpLpSolveInfo)]),
        "gp_preprocess_batch": (C.c_int, [C.POINTER(GpDeviceLp), C.c_int32, C.POINTER(GpParams),
                                          C.POINTER(GpBatchSummary)]),
        "gp_postprocess_batch": (C.c_int, [vp, vp, vp, C.POINTER(GpBatchSummary), vp, vp, vp]),
        "gp_batch_instance": (C.c_int, [C.POINTER(GpBatchSummary), C.c_int32, C.POINTER(GpDeviceLp)]),
        "gp_free_batch_summary": (None, [C.POINTER(GpBatchSummary)]),
        "gp_get_tuning": (C.c_int, [C.POINTER(GpTuning)]),
        "gp_set_tuning": (C.c_int, [C.POINTER(GpTuning)]),
        "gp_autotune": (C.c_int, [C.POINTER(GpDeviceLp), C.POINTER(GpParams), C.c_int32, C.POINTER(GpTuning),
                                  C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "gp_rule_group": (C.c_int, [C.c_int32]),
        "gp_schedule_features": (None, [C.POINTER(GpRoundTrace), C.c_int32, C.c_int32, C.c_int64, C.c_int64,
                                        C.c_int32, C.c_int32, C.c_int32, C.c_int32, vp]),
        "gp_mlp_forward": (C.c_int, [vp, vp, C.c_int32, vp]),
        "gp_mlp_loss_grad": (C.c_int, [vp, vp, vp, vp, C.c_int32, vp, vp]),
        "gp_mlp_adam_step": (C.c_int, [vp, vp, vp, vp, C.c_int32, C.c_float, C.c_float, C.c_float, C.c_float]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class GpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"gpupre error {code}: {msg}")
        self.code = code


def check(code: int):
    if code != 0:
        raise GpError(code, load().gp_last_error().decode(errors="replace"))


def device_count() -> int:
    n = C.c_int32(0)
    check(load().gp_device_count(C.byref(n)))
    return int(n.value)


def default_params(**overrides) -> GpParams:
    p = GpParams()
    load().gp_default_params(C.byref(p))
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def parity_params(**overrides) -> GpParams:
    """Reference defaults with folding and structure-specific rules off."""
    base = dict(enable_folding=0, enable_structure_specific_rules=0)
    base.update(overrides)
    return default_params(**base)


# ---------------------------------------------------------------- device memory
class DeviceArray:
    """A device allocation owned by Python (gp_malloc / gp_free)."""

    def __init__(self, nbytes: int):
        self.ptr = C.c_void_p()
        self.nbytes = int(nbytes)
        check(load().gp_malloc(C.byref(self.ptr), max(1, self.nbytes)))

    @staticmethod
    def from_host(a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = DeviceArray(a.nbytes)
        if a.nbytes:
            check(load().gp_memcpy_h2d(d.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes))
        return d

    def free(self):
        if self.ptr:
            check(load().gp_free(self.ptr))
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def to_host(ptr, count: int, dtype) -> np.ndarray:
    out = np.empty(int(count), dtype=dtype)
    if count:
        check(load().gp_memcpy_d2h(out.ctypes.data_as(C.c_void_p), ptr, out.nbytes))
    return out


def to_device(ptr, a: np.ndarray):
    a = np.ascontiguousarray(a)
    if a.nbytes:
        check(load().gp_memcpy_h2d(ptr, a.ctypes.data_as(C.c_void_p), a.nbytes))


class DeviceLP:
    """A host LP uploaded to HBM (A by copy, A^T built on the device)."""

    def __init__(self, lp: HostLP):
        self.lp = GpDeviceLp()
        self.m, self.n = lp.m, lp.n
        check(load().gp_lp_upload(lp.m, lp.n, lp.nnz, _ptr(lp.row_ptr), _ptr(lp.col_idx), _ptr(lp.vals),
                                  _ptr(lp.c), _ptr(lp.AL), _ptr(lp.AU), _ptr(lp.l), _ptr(lp.u),
                                  float(lp.obj_constant), C.byref(self.lp)))

    @classmethod
    def from_mps(cls, model: MpsModel) -> "DeviceLP":
        """gp_mps_upload: duplicates summed and A, A^T assembled on the device."""
        self = cls.__new__(cls)
        self.lp = GpDeviceLp()
        self.m, self.n = model.m, model.n
        check(load().gp_mps_upload(C.byref(model.raw), C.byref(self.lp)))
        return self

    @classmethod
    def from_entries(cls, m, n, rows, cols, vals, c, AL, AU, l, u, obj_constant=0.0) -> "DeviceLP":
        """gp_lp_upload_entries on host (row, column, value) entries in any order."""
        self = cls.__new__(cls)
        self.lp = GpDeviceLp()
        self.m, self.n = m, n
        rows, cols = np.ascontiguousarray(rows, np.int32), np.ascontiguousarray(cols, np.int32)
        vals = np.ascontiguousarray(vals, np.float64)
        vecs = [np.ascontiguousarray(v, np.float64) for v in (c, AL, AU, l, u)]
        check(load().gp_lp_upload_entries(m, n, len(vals), _ptr(rows), _ptr(cols), _ptr(vals),
                                          *[_ptr(v) for v in vecs], float(obj_constant), C.byref(self.lp)))
        return self

    def free(self):
        if self.lp.A.row_ptr:
            load().gp_lp_free(C.byref(self.lp))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class MpsModel:
    """An MPS file read by gp_mps_parse / gp_mps_read (host memory of the
    library): dimensions, vectors and the matrix entries in file order."""

    def __init__(self, raw: GpMpsModel):
        self.raw = raw
        self.m, self.n, self.entries = raw.rows, raw.cols, raw.entries
        self.name = C.string_at(raw.name) if raw.name else b""
        self.obj_constant = float(raw.obj_constant)

    @staticmethod
    def _host(ptr, count, ctype, dtype):
        if count == 0:
            return np.zeros(0, dtype)
        return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ctype)), (count,)).copy()

    def arrays(self) -> dict:
        r, i32, f64 = self.raw, (C.c_int32, np.int32), (C.c_double, np.float64)
        return {"entry_row": self._host(r.entry_row, self.entries, *i32),
                "entry_col": self._host(r.entry_col, self.entries, *i32),
                "entry_val": self._host(r.entry_val, self.entries, *f64),
                "c": self._host(r.c, self.n, *f64), "l": self._host(r.l, self.n, *f64),
                "u": self._host(r.u, self.n, *f64), "AL": self._host(r.AL, self.m, *f64),
                "AU": self._host(r.AU, self.m, *f64)}

    def upload(self) -> "DeviceLP":
        return DeviceLP.from_mps(self)

    def free(self):
        load().gp_mps_free(C.byref(self.raw))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()


def write_mps(lp: HostLP, name: str = "lp") -> bytes:
    """Free-format MPS text of a host LP (rows R<i>, columns X<j>, numbers in
    repr form; a column without entries or cost gets a zero-cost card so
    that it exists for a reader).  Rows are E / L / G with RANGES for the
    two-sided ones; a free row is an L row with an infinite right-hand side."""
    import io

    out = io.StringIO()
    AL, AU = np.asarray(lp.AL), np.asarray(lp.AU)
    kind = np.where(AL == AU, "E", np.where(np.isfinite(AU), "L", "G"))
    out.write(f"NAME {name}\nROWS\n N obj\n")
    out.writelines(f" {k} R{i}\n" for i, k in enumerate(kind))
    out.write("COLUMNS\n")
    rows = np.repeat(np.arange(lp.m, dtype=np.int64), np.diff(np.asarray(lp.row_ptr, np.int64)))
    order = np.lexsort((rows, np.asarray(lp.col_idx)))
    cols = np.asarray(lp.col_idx)[order]
    ptr = np.searchsorted(cols, np.arange(lp.n + 1))
    r_sorted, v_sorted = rows[order], np.asarray(lp.vals)[order]
    c = np.asarray(lp.c)
    for j in range(lp.n):
        if c[j] != 0.0 or ptr[j] == ptr[j + 1]:
            out.write(f" X{j} obj {float(c[j])!r}\n")
        out.writelines(f" X{j} R{r} {float(v)!r}\n" for r, v in zip(r_sorted[ptr[j]:ptr[j + 1]],
                                                                   v_sorted[ptr[j]:ptr[j + 1]]))
    out.write("RHS\n")
    for i in range(lp.m):
        side = AU[i] if kind[i] in ("E", "L") else AL[i]
        if side != 0.0:
            out.write(f" rhs R{i} {float(side)!r}\n")
    two = np.flatnonzero((kind == "L") & np.isfinite(AL))
    if two.size:
        out.write("RANGES\n")
        out.writelines(f" rng R{i} {float(AU[i] - AL[i])!r}\n" for i in two)
    out.write("BOUNDS\n")
    l, u = np.asarray(lp.l), np.asarray(lp.u)
    for j in range(lp.n):
        if np.isfinite(l[j]) and l[j] == u[j]:
            out.write(f" FX bnd X{j} {float(l[j])!r}\n")
            continue
        if not np.isfinite(l[j]) and not np.isfinite(u[j]):
            out.write(f" FR bnd X{j}\n")
            continue
        if not np.isfinite(l[j]):
            out.write(f" MI bnd X{j}\n")
        elif l[j] != 0.0:
            out.write(f" LO bnd X{j} {float(l[j])!r}\n")
        if np.isfinite(u[j]):
            out.write(f" UP bnd X{j} {float(u[j])!r}\n")
    out.write("ENDATA\n")
    return out.getvalue().encode()


def parse_mps(text) -> MpsModel:
    """gp_mps_parse on an in-memory MPS text (bytes or str)."""
    if isinstance(text, str):
        text = text.encode("latin-1")
    raw = GpMpsModel()
    check(load().gp_mps_parse(text, len(text), C.byref(raw)))
    return MpsModel(raw)


def read_mps(path) -> MpsModel:
    """gp_mps_read on a file."""
    raw = GpMpsModel()
    check(load().gp_mps_read(os.fsencode(path), C.byref(raw)))
    return MpsModel(raw)


class Fold:
    """gp_fold: an LP folded by colour refinement (device memory owned here)."""

    def __init__(self, lp, tolerance: float = 1e-8):
        d = lp.lp if isinstance(lp, DeviceLP) else lp
        self.r = GpFoldResult()
        check(load().gp_fold(C.byref(d), float(tolerance), C.byref(self.r)))
        self.m, self.n = d.A.rows, d.A.cols

    @property
    def applied(self) -> bool:
        return bool(self.r.applied)

    def classes(self):
        """(row class, column class) host arrays, 1-based (applied folds only)."""
        return to_host(self.r.row_color_id, self.m, np.int32), to_host(self.r.col_color_id, self.n, np.int32)

    def scales(self):
        return to_host(self.r.row_scale, self.m, np.float64), to_host(self.r.col_scale, self.n, np.float64)

    def folded(self) -> HostLP:
        return device_lp_to_host(self.r.folded_lp)

    def unfold(self, x_red, y_red, z_red):
        """Reduced host vectors in, original-order host vectors out."""
        dx, dy, dz = (DeviceArray.from_host(np.asarray(v, np.float64)) for v in (x_red, y_red, z_red))
        x, y, z = np.empty(self.n), np.empty(self.m), np.empty(self.n)
        check(load().gp_unfold(C.byref(self.r), dx.ptr, dy.ptr, dz.ptr, _ptr(x), _ptr(y), _ptr(z)))
        return x, y, z

    def free(self):
        load().gp_free_fold(C.byref(self.r))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()


def lp_solve_params(**overrides) -> GpLpSolveParams:
    p = GpLpSolveParams()
    load().gp_lp_solve_default_params(C.byref(p))
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def solve_lp(lp, params: GpLpSolveParams | None = None, **overrides):
    """gp_lp_solve on a device LP (DeviceLP or GpDeviceLp): host x, y, z and the info struct."""
    d = lp.lp if isinstance(lp, DeviceLP) else lp
    m, n = d.A.rows, d.A.cols
    p = params if params is not None else lp_solve_params(**overrides)
    dx, dy, dz = DeviceArray(8 * n), DeviceArray(8 * m), DeviceArray(8 * n)
    info = GpLpSolveInfo()
    check(load().gp_lp_solve(C.byref(d), C.byref(p), dx.ptr, dy.ptr, dz.ptr, C.byref(info)))
    return to_host(dx.ptr, n, np.float64), to_host(dy.ptr, m, np.float64), to_host(dz.ptr, n, np.float64), info


def device_lp_to_host(d: GpDeviceLp) -> HostLP:
    m, n, nnz = d.A.rows, d.A.cols, d.A.nnz
    h = HostLP(m, n, to_host(d.A.row_ptr, m + 1, NP_OFFSET), to_host(d.A.col_idx, nnz, np.int32),
               to_host(d.A.vals, nnz, np.float64), to_host(d.c, n, np.float64), to_host(d.l, n, np.float64),
               to_host(d.u, n, np.float64), to_host(d.AL, m, np.float64), to_host(d.AU, m, np.float64),
               float(d.obj_constant))
    h.at = (to_host(d.AT.row_ptr, n + 1, NP_OFFSET), to_host(d.AT.col_idx, d.AT.nnz, np.int32),
            to_host(d.AT.vals, d.AT.nnz, np.float64))
    return h


# ---------------------------------------------------------------- public pipeline
@dataclass
class Record:
    m0: int
    n0: int
    m1: int
    n1: int
    row_red2org: np.ndarray
    col_red2org: np.ndarray
    fixed_idx: np.ndarray
    fixed_val: np.ndarray
    types: np.ndarray
    index_starts: np.ndarray
    value_starts: np.ndarray
    dual_modes: np.ndarray
    indices: np.ndarray
    vals: np.ndarray


class Summary:
    """gp_summary wrapper (FoldingPipelineSummary); call free() when done."""

    def __init__(self):
        self.s = GpSummary()

    def free(self):
        load().gp_free_summary(C.byref(self.s))

    def __getattr__(self, name):
        return getattr(self.s, name)

    def record(self) -> Record:
        sz = GpRecordSizes()
        check(load().gp_record_get_sizes(C.byref(self.s), C.byref(sz)))
        a = lambda k, dt: np.empty(int(k), dtype=dt)
        rec = Record(sz.m0, sz.n0, sz.m1, sz.n1, a(sz.m1, np.int32), a(sz.n1, np.int32),
                     a(sz.fixed_count, np.int32), a(sz.fixed_count, np.float64),
                     a(sz.record_count, np.int32), a(sz.record_count + 1, NP_OFFSET),
                     a(sz.record_count + 1, NP_OFFSET), a(sz.record_count, np.uint8),
                     a(sz.index_count, np.int32), a(sz.value_count, np.float64))
        check(load().gp_record_copy(C.byref(self.s), _ptr(rec.row_red2org), _ptr(rec.col_red2org),
                                    _ptr(rec.fixed_idx), _ptr(rec.fixed_val), _ptr(rec.types),
                                    _ptr(rec.index_starts), _ptr(rec.value_starts), _ptr(rec.dual_modes),
                                    _ptr(rec.indices), _ptr(rec.vals)))
        return rec

    def reduced(self) -> HostLP:
        return device_lp_to_host(self.s.reduced_lp)

    def trace(self):
        return [GpRoundTrace.from_buffer_copy(self.s.trace[i]) for i in range(self.s.trace_count)]

    def kernel_stats(self):
        return {name: dict(ms=self.s.kernel_ms[k], launches=int(self.s.kernel_launches[k]),
                           bytes=self.s.kernel_bytes[k]) for k, name in enumerate(KERNEL_FAMILIES)}


def preprocess(dlp: DeviceLP, params: GpParams | None = None) -> Summary:
    if params is None:
        params = default_params()
    s = Summary()
    check(load().gp_preprocess(C.byref(dlp.lp), C.byref(params), C.byref(s.s)))
    return s


def postprocess(summary: Summary, x: np.ndarray, y: np.ndarray, z: np.ndarray, original: DeviceLP | None):
    """x, y, z in reduced order (host arrays, uploaded here); returns original-order host arrays."""
    rec = GpRecordSizes()
    check(load().gp_record_get_sizes(C.byref(summary.s), C.byref(rec)))
    dx, dy, dz = (DeviceArray.from_host(np.asarray(v, dtype=np.float64)) for v in (x, y, z))
    # original order: the folding stage's input when it folded, else the presolve's
    f = summary.s.folding
    m0, n0 = (f.original_rows, f.original_cols) if summary.s.folding_applied else (rec.m0, rec.n0)
    xo = np.empty(n0)
    yo = np.empty(m0)
    zo = np.empty(n0)
    check(load().gp_postprocess(dx.ptr, dy.ptr, dz.ptr, C.byref(summary.s),
                                C.byref(original.lp) if original is not None else None,
                                _ptr(xo), _ptr(yo), _ptr(zo)))
    return xo, yo, zo


# ---------------------------------------------------------------- postsolve of a built record
class GpAbsValueStep(C.Structure):
    _fields_ = [("split_count", C.c_int32), ("split_t_col", C.c_void_p), ("split_e_col", C.c_void_p),
                ("split_rho", C.c_void_p), ("pair_count", C.c_int32), ("pair_bound_col", C.c_void_p),
                ("pair_free_col", C.c_void_p), ("linked_count", C.c_int32), ("linked_slack_col", C.c_void_p),
                ("linked_t_col", C.c_void_p), ("linked_factor", C.c_void_p), ("max_count", C.c_int32),
                ("max_slack_col", C.c_void_p), ("max_starts", C.c_void_p), ("max_t_cols", C.c_void_p),
                ("max_factors", C.c_void_p)]


class GpVarPairStep(C.Structure):
    _fields_ = [("count", C.c_int32), ("plus_cols", C.c_void_p), ("minus_cols", C.c_void_p),
                ("root_plus_cols", C.c_void_p), ("root_minus_cols", C.c_void_p), ("plus_shift", C.c_void_p),
                ("minus_shift", C.c_void_p)]


class GpCheckpoint(C.Structure):
    _fields_ = [("kind", C.c_int32), ("payload_index", OFFSET), ("tape_position", OFFSET)]


class GpHostRecord(C.Structure):
    _fields_ = [("m0", C.c_int32), ("n0", C.c_int32), ("m1", C.c_int32), ("n1", C.c_int32),
                ("row_red2org", C.c_void_p), ("col_red2org", C.c_void_p), ("fixed_count", C.c_int32),
                ("fixed_idx", C.c_void_p), ("fixed_val", C.c_void_p), ("record_count", OFFSET),
                ("types", C.c_void_p), ("index_starts", C.c_void_p), ("value_starts", C.c_void_p),
                ("dual_modes", C.c_void_p), ("index_count", OFFSET), ("indices", C.c_void_p),
                ("value_count", OFFSET), ("vals", C.c_void_p), ("abs_value_count", C.c_int32),
                ("abs_value", C.POINTER(GpAbsValueStep)), ("var_pair_count", C.c_int32),
                ("var_pair", C.POINTER(GpVarPairStep)), ("checkpoint_count", C.c_int32),
                ("checkpoints", C.POINTER(GpCheckpoint))]


class HostRecord:
    """PresolveRecordGpu as host arrays (gp_host_record), kept alive here.

    maps / fixed / tape arguments are array-likes; abs_value is a list of
    dicts with optional keys splits [(t, e, rho)], outer_pairs [(bound, free)],
    linked_slacks [(slack, t, factor)], max_slacks [(slack, [t...], [f...])];
    var_pair is a list of dicts with plus_cols, minus_cols, root_plus_cols,
    root_minus_cols, plus_shift, minus_shift; checkpoints are
    (kind, payload_index, tape_position) triples."""

    def __init__(self, m0, n0, m1, n1, row_red2org=(), col_red2org=(), fixed_idx=(), fixed_val=(), types=(),
                 index_starts=(0,), value_starts=(0,), dual_modes=(), indices=(), vals=(), abs_value=(),
                 var_pair=(), checkpoints=()):
        keep = self._keep = []

        def arr(v, dt):
            a = np.ascontiguousarray(np.asarray(v, dtype=dt))
            keep.append(a)
            return a.ctypes.data_as(C.c_void_p) if a.size else None

        r = self.rec = GpHostRecord()
        r.m0, r.n0, r.m1, r.n1 = int(m0), int(n0), int(m1), int(n1)
        r.row_red2org, r.col_red2org = arr(row_red2org, np.int32), arr(col_red2org, np.int32)
        r.fixed_count = len(fixed_idx)
        r.fixed_idx, r.fixed_val = arr(fixed_idx, np.int32), arr(fixed_val, np.float64)
        r.record_count = len(types)
        r.types = arr(types, np.int32)
        r.index_starts, r.value_starts = arr(index_starts, NP_OFFSET), arr(value_starts, NP_OFFSET)
        r.dual_modes = arr(dual_modes, np.uint8)
        r.index_count, r.value_count = len(indices), len(vals)
        r.indices, r.vals = arr(indices, np.int32), arr(vals, np.float64)
        steps = (GpAbsValueStep * max(1, len(abs_value)))()
        for k, st in enumerate(abs_value):
            sp = list(st.get("splits", ()))
            op = list(st.get("outer_pairs", ()))
            ls = list(st.get("linked_slacks", ()))
            ms = list(st.get("max_slacks", ()))
            o = steps[k]
            o.split_count = len(sp)
            o.split_t_col = arr([a[0] for a in sp], np.int32)
            o.split_e_col = arr([a[1] for a in sp], np.int32)
            o.split_rho = arr([a[2] for a in sp], np.float64)
            o.pair_count = len(op)
            o.pair_bound_col = arr([a[0] for a in op], np.int32)
            o.pair_free_col = arr([a[1] for a in op], np.int32)
            o.linked_count = len(ls)
            o.linked_slack_col = arr([a[0] for a in ls], np.int32)
            o.linked_t_col = arr([a[1] for a in ls], np.int32)
            o.linked_factor = arr([a[2] for a in ls], np.float64)
            o.max_count = len(ms)
            o.max_slack_col = arr([a[0] for a in ms], np.int32)
            o.max_starts = arr(np.concatenate([[0], np.cumsum([len(a[1]) for a in ms])]) if ms else [0], NP_OFFSET)
            o.max_t_cols = arr([c for a in ms for c in a[1]], np.int32)
            o.max_factors = arr([f for a in ms for f in a[2]], np.float64)
        keep.append(steps)
        r.abs_value_count = len(abs_value)
        r.abs_value = C.cast(steps, C.POINTER(GpAbsValueStep))
        pairs = (GpVarPairStep * max(1, len(var_pair)))()
        for k, st in enumerate(var_pair):
            o = pairs[k]
            o.count = len(st["plus_cols"])
            o.plus_cols, o.minus_cols = arr(st["plus_cols"], np.int32), arr(st["minus_cols"], np.int32)
            o.root_plus_cols = arr(st["root_plus_cols"], np.int32)
            o.root_minus_cols = arr(st["root_minus_cols"], np.int32)
            o.plus_shift, o.minus_shift = arr(st["plus_shift"], np.float64), arr(st["minus_shift"], np.float64)
        keep.append(pairs)
        r.var_pair_count = len(var_pair)
        r.var_pair = C.cast(pairs, C.POINTER(GpVarPairStep))
        cps = (GpCheckpoint * max(1, len(checkpoints)))()
        for k, (kind, payload, pos) in enumerate(checkpoints):
            cps[k].kind, cps[k].payload_index, cps[k].tape_position = int(kind), int(payload), int(pos)
        keep.append(cps)
        r.checkpoint_count = len(checkpoints)
        r.checkpoints = C.cast(cps, C.POINTER(GpCheckpoint))

    @staticmethod
    def from_record(rec: "Record") -> "HostRecord":
        return HostRecord(rec.m0, rec.n0, rec.m1, rec.n1, rec.row_red2org, rec.col_red2org, rec.fixed_idx,
                          rec.fixed_val, rec.types, rec.index_starts, rec.value_starts, rec.dual_modes,
                          rec.indices, rec.vals)


def postsolve_record(rec: HostRecord, x, y, z, original: GpDeviceLp | None = None, original_l=None):
    """gp_postsolve_record: x, y, z host arrays in reduced order or None (a
    null device pointer); `original` a device LP struct (fields may be null);
    returns original-order host x, y, z."""
    r = rec.rec
    dev = [DeviceArray.from_host(np.asarray(v, dtype=np.float64)) if v is not None else None for v in (x, y, z)]
    ptrs = [d.ptr if d is not None else None for d in dev]
    xo, yo, zo = np.empty(r.n0), np.empty(r.m0), np.empty(r.n0)
    check(load().gp_postsolve_record(ptrs[0], ptrs[1], ptrs[2], C.byref(r),
                                     C.byref(original) if original is not None else None,
                                     original_l, _ptr(xo), _ptr(yo), _ptr(zo)))
    return xo, yo, zo


# ---------------------------------------------------------------- batched presolve
class BatchSummary:
    """Owner of a gp_batch_summary (presolve of several device LPs as one batch)."""

    def __init__(self):
        self.s = GpBatchSummary()

    def free(self):
        load().gp_free_batch_summary(C.byref(self.s))

    @property
    def count(self) -> int:
        return self.s.count

    def offsets(self, name) -> np.ndarray:
        return np.ctypeslib.as_array(getattr(self.s, name), shape=(self.s.count + 1,)).copy()

    def obj_constants(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.s.obj_constant, shape=(self.s.count,)).copy()

    def status(self) -> np.ndarray:
        """Per instance: 0 presolved in the batch, 1 infeasible, 2 unbounded (retired alone)."""
        return np.ctypeslib.as_array(self.s.status, shape=(self.s.count,)).copy()

    @property
    def restarts(self) -> int:
        return self.s.restarts

    def whole(self) -> "Summary":
        """The direct-sum summary as a Summary view (owned by this batch)."""
        v = Summary()
        v.s = self.s.whole
        v.free = lambda: None
        return v

    def instance(self, k: int) -> HostLP:
        """Reduced instance k, extracted on the device and copied to the host."""
        d = GpDeviceLp()
        check(load().gp_batch_instance(C.byref(self.s), int(k), C.byref(d)))
        try:
            return device_lp_to_host(d)
        finally:
            load().gp_lp_free(C.byref(d))

    def postprocess(self, x, y, z):
        """x, y, z over the reduced direct sum (host arrays, uploaded here); returns
        host arrays over the original direct sum."""
        m0, n0 = int(self.offsets("row_offsets")[-1]), int(self.offsets("col_offsets")[-1])
        dx, dy, dz = (DeviceArray.from_host(np.asarray(v, dtype=np.float64)) for v in (x, y, z))
        xo, yo, zo = np.empty(n0), np.empty(m0), np.empty(n0)
        check(load().gp_postprocess_batch(dx.ptr, dy.ptr, dz.ptr, C.byref(self.s), _ptr(xo), _ptr(yo), _ptr(zo)))
        for a in (dx, dy, dz):
            a.free()
        return xo, yo, zo


def autotune(lp: GpDeviceLp, params: GpParams | None = None, reps: int = 1):
    """(best GpTuning, tuned ms, default ms) measured on a device LP."""
    best = GpTuning()
    t, d = C.c_double(), C.c_double()
    check(load().gp_autotune(C.byref(lp), C.byref(params or parity_params()), int(reps), C.byref(best),
                             C.byref(t), C.byref(d)))
    return best, t.value, d.value


def set_tuning(t: GpTuning | None = None):
    check(load().gp_set_tuning(C.byref(t or GpTuning())))


def preprocess_batch(dlps, params: GpParams | None = None) -> BatchSummary:
    arr = (GpDeviceLp * len(dlps))(*[d.lp for d in dlps])
    out = BatchSummary()
    check(load().gp_preprocess_batch(arr, len(dlps), C.byref(params or parity_params()), C.byref(out.s)))
    return out


def direct_sum(lps) -> HostLP:
    """Host block-diagonal direct sum of host LPs (the batch's assembled LP)."""
    rp = [np.zeros(1, dtype=NP_OFFSET)]
    ci, va = [], []
    nnz = 0
    col = 0
    for lp in lps:
        rp.append(np.asarray(lp.row_ptr[1:], dtype=np.int64) + nnz)
        ci.append(np.asarray(lp.col_idx, dtype=np.int64) + col)
        va.append(lp.vals)
        nnz += lp.nnz
        col += lp.n
    cat = np.concatenate
    return HostLP(sum(lp.m for lp in lps), col, cat(rp).astype(NP_OFFSET), cat(ci).astype(np.int32), cat(va),
                  cat([lp.c for lp in lps]), cat([lp.l for lp in lps]), cat([lp.u for lp in lps]),
                  cat([lp.AL for lp in lps]), cat([lp.AU for lp in lps]), float(sum(lp.obj_constant for lp in lps)))


# ---------------------------------------------------------------- generator
class LpgenSpec(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("seed", C.c_uint64), ("degree_mode", C.c_int32),
                ("mean_degree", C.c_double), ("zipf_s", C.c_double), ("zipf_cap", C.c_int32),
                ("frac_equality", C.c_double), ("frac_free", C.c_double),
                ("frac_singleton_rows", C.c_double), ("frac_doubleton_rows", C.c_double),
                ("frac_parallel_rows", C.c_double), ("frac_duplicate_cols", C.c_double),
                ("frac_fixed_cols", C.c_double), ("frac_empty_cols", C.c_double), ("integral", C.c_int32)]


class LpgenLp(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("nnz", C.c_int32), ("row_ptr", C.c_void_p),
                ("col_idx", C.c_void_p), ("vals", C.c_void_p), ("c", C.c_void_p), ("l", C.c_void_p),
                ("u", C.c_void_p), ("AL", C.c_void_p), ("AU", C.c_void_p), ("x0", C.c_void_p),
                ("y0", C.c_void_p), ("obj_constant", C.c_double), ("zipf_scale", C.c_double)]


def _load_lpgen():
    global _lpgen
    if _lpgen is None:
        if not os.path.exists(LPGEN_PATH):
            raise RuntimeError(f"{LPGEN_PATH} is not built; run __graft_entry__.build()")
        g = C.CDLL(LPGEN_PATH)
        g.lpgen_default_spec.argtypes = [C.POINTER(LpgenSpec), C.c_int32, C.c_int32, C.c_uint64]
        g.lpgen_default_spec.restype = None
        g.lpgen_generate.argtypes = [C.POINTER(LpgenSpec), C.POINTER(LpgenLp)]
        g.lpgen_generate.restype = C.c_int
        g.lpgen_free.argtypes = [C.POINTER(LpgenLp)]
        g.lpgen_free.restype = None
        g.lpgen_qap.argtypes = [C.c_int32, C.c_uint64, C.POINTER(LpgenLp)]
        g.lpgen_qap.restype = C.c_int
        _lpgen = g
    return _lpgen


# Workload definitions (sizes of the synthetic presolve benchmarks).
WORKLOADS = {
    # 2,000 x 3,000, ~6 nnz per column, uniform rows
    "small": dict(m=2000, n=3000, degree_mode=0, mean_degree=6.0),
    # 200,000 x 300,000, ~2M nnz, Zipf(1.5) column degree capped at 500, 1% empty columns
    "medium": dict(m=200_000, n=300_000, degree_mode=1, mean_degree=2_000_000 / 300_000, frac_empty_cols=0.01),
    # 5M x 8M, ~60M nnz, same recipe with integer data (exact arithmetic: the
    # presolve reaches its fixpoint instead of stopping on a tolerance artefact)
    "large": dict(m=5_000_000, n=8_000_000, degree_mode=1, mean_degree=60_000_000 / 8_000_000,
                  frac_empty_cols=0.01, integral=1),
    # the large recipe with FP64 N(0,1) values as in the config text; the
    # reference's tolerance policy ends it as infeasible after two rounds
    "large_fp64": dict(m=5_000_000, n=8_000_000, degree_mode=1, mean_degree=60_000_000 / 8_000_000,
                       frac_empty_cols=0.01),
    # the large recipe at 1/10 scale (parity cases for the bench workload)
    "large10": dict(m=500_000, n=800_000, degree_mode=1, mean_degree=60_000_000 / 8_000_000,
                    frac_empty_cols=0.01, integral=1),
}


def generate(m: int, n: int, seed: int = 0, **spec_overrides) -> HostLP:
    g = _load_lpgen()
    spec = LpgenSpec()
    g.lpgen_default_spec(C.byref(spec), int(m), int(n), int(seed))
    for k, v in spec_overrides.items():
        setattr(spec, k, v)
    raw = LpgenLp()
    if g.lpgen_generate(C.byref(spec), C.byref(raw)) != 0:
        raise RuntimeError("lpgen_generate failed")
    return _take_lpgen(g, raw)


def generate_qap(n: int, seed: int = 0) -> HostLP:
    """Adams-Johnson LP linearisation of a seeded n x n QAP (rows 2n^4+2n^3+n^2+2n,
    columns n^4+n^2), in the row and column order of scripts/generate_qap20.jl:25-109."""
    g = _load_lpgen()
    raw = LpgenLp()
    if g.lpgen_qap(int(n), int(seed), C.byref(raw)) != 0:
        raise RuntimeError(f"lpgen_qap failed for n={n}")
    return _take_lpgen(g, raw)


def _take_lpgen(g, raw) -> HostLP:
    def arr(p, k, dt):
        if k == 0:
            return np.empty(0, dtype=dt)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(np.ctypeslib.as_ctypes_type(dt))), shape=(k,)).copy()

    lp = HostLP(raw.m, raw.n, arr(raw.row_ptr, raw.m + 1, np.int32), arr(raw.col_idx, raw.nnz, np.int32),
                arr(raw.vals, raw.nnz, np.float64), arr(raw.c, raw.n, np.float64), arr(raw.l, raw.n, np.float64),
                arr(raw.u, raw.n, np.float64), arr(raw.AL, raw.m, np.float64), arr(raw.AU, raw.m, np.float64),
                float(raw.obj_constant), arr(raw.x0, raw.n, np.float64), arr(raw.y0, raw.m, np.float64))
    g.lpgen_free(C.byref(raw))
    return lp


def batch_workload(count: int = 512, seed: int = 0, min_rows: int = 1000, max_rows: int = 100_000,
                   threads: int = 8, integral: int = 0) -> list:
    """The batched-presolve workload: `count` independent LPs of the synthetic
    generator (uniform column degree 6, planted structure), rows log-uniform in
    [min_rows, max_rows], columns 1.5 x rows; instance k uses seed seed*100003+k.
    integral=1 draws integer coefficients, bounds and planted point."""
    import concurrent.futures as cf

    rng = np.random.default_rng(seed)
    rows = np.exp(rng.uniform(np.log(min_rows), np.log(max_rows), size=count)).astype(np.int64)

    def one(k):
        m = int(rows[k])
        return generate(m, m * 3 // 2, seed * 100003 + k, degree_mode=0, mean_degree=6.0, integral=integral)

    with cf.ThreadPoolExecutor(max_workers=threads) as pool:
        return list(pool.map(one, range(count)))


def generate_workload(name: str, seed: int = 0) -> HostLP:
    w = dict(WORKLOADS[name])
    m, n = w.pop("m"), w.pop("n")
    return generate(m, n, seed, **w)
