"""The C-ABI library builds, loads and exports every symbol include/gpupre.h declares."""
import ctypes as C
import os
import re

import gpupre as G

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(REPO, "include", "gpupre.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gp_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(G.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_abi_version_and_defaults():
    lib = G.load()
    assert lib.gp_abi_version() == 1
    p = G.default_params()
    assert p.struct_size == C.sizeof(G.GpParams)
    # reference defaults (presolve_structs.hpp:28-158)
    assert p.max_iters == 512 and p.feasibility_tol == 1e-6 and p.bound_tol == 1e-6 and p.zero_tol == 1e-10
    assert p.doubleton_equations_min_selected_per_batch == 256
    assert p.dual_fix_max_nonzero_cost_col_degree == 128
    assert p.bound_strengthening_bound_only_nnz_round_budget == 8000000
    assert p.record_postsolve_tape == 1 and p.record_postsolve_tape_cpu == 0


def test_struct_sizes_match_c_layout():
    # gp_csr: 3 x int32 + 3 pointers; gp_device_lp: 2 csr + 5 pointers + double
    assert C.sizeof(G.GpCsr) == 40
    assert C.sizeof(G.GpDeviceLp) == 2 * 40 + 5 * 8 + 8


def test_errors_are_status_codes_without_device():
    lib = G.load()
    s = G.GpSummary()
    p = G.default_params()
    p.struct_size = 3
    rc = lib.gp_preprocess(C.byref(G.GpDeviceLp()), C.byref(p), C.byref(s))
    assert rc == 1  # GP_ERR_INVALID_ARGUMENT
    assert b"struct_size" in lib.gp_last_error()
    lib.gp_free_summary(C.byref(s))
    lib.gp_free_summary(C.byref(s))  # idempotent
    lib.gp_free_summary(None)


def test_lp_solver_defaults_and_layouts():
    p = G.lp_solve_params()
    # HPR-LP's defaults (third_party/hprlp/include/api/structs.h:256-291): stop_tol 1e-6, time_limit 1000 s
    assert p.struct_size == C.sizeof(G.GpLpSolveParams) == 96
    assert p.tolerance == 1e-6 and p.time_limit == 1000.0 and p.check_every == 64
    assert p.ruiz_iterations == 10 and p.pock_chambolle == 1 and p.bound_cost_scaling == 1
    assert C.sizeof(G.GpLpSolveInfo) == 16 + 9 * 8
    assert C.sizeof(G.GpMpsModel) == 4 + 4 + 8 + 8 * 8 + 8 + 8
