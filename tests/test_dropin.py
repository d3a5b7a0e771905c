"""The C++ front end with the reference's API names (include/gpu_preprocessor/
gpu_preprocessor.hpp): it compiles as plain host C++ (layout checks against
the C ABI are static_asserts), and on a GPU the example caller
(examples/presolve_example.cpp) runs preprocess / postprocess / free_summary
on an LP that presolve solves completely and checks the recovered point's
KKT conditions; the MPS tool (examples/mps_presolve.cpp) reads a file,
uploads it and presolves it with the same outcome as the Python path."""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "gpu-preprocessor-ijoc_amd")


def test_shim_compiles_as_host_cpp(tmp_path):
    src = tmp_path / "use_shim.cpp"
    src.write_text(
        '#include "gpu_preprocessor/gpu_preprocessor.hpp"\n'
        "int main() {\n"
        "  gpu_preprocessor::Parameters p;\n"
        "  p.enable_folding = false;\n"
        "  gpu_preprocessor::DeviceLp lp;\n"
        "  lp.A = {2, 2, 2, nullptr, nullptr, nullptr};\n"
        "  gpu_preprocessor::Summary s;\n"
        "  gpu_preprocessor::free_summary(s);\n"
        "  gpu_preprocessor::tools::MpsModel model;\n"
        "  gpu_preprocessor::tools::DeviceLpOwner owner;\n"
        "  if (model.num_rows() != 0 || owner.lp.A.rowPtr) return 2;\n"
        "  return p.max_iters == 512 && lp.A.rows == 2 ? 0 : 1;\n"
        "}\n")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(REPO, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


ADAPTER = r"""
// A solver adapter written against the drop-in names the way the reference's
// callers use them (solver_interface/adapters/gpu.cpp:700-780,
// cpp/tools/gpu_preprocessor_mps.cpp:30-62): parameters set field by field,
// preprocess, folding and presolve timers and dimensions, the record's
// objective constant and exported fixed columns, postprocess, free_summary.
#include <cmath>
#include <iostream>
#include "gpu_preprocessor/gpu_preprocessor.hpp"

namespace gp = gpu_preprocessor;

struct Handle {
  gp::DeviceLp original_lp;
  gp::Summary folding_summary;
  double presolve_seconds = 0.0;
};

bool run_adapter(Handle& h, bool folding, bool presolve, bool new_rules, double time_limit) {
  gp::Parameters p;
  p.verbose = false;
  p.enable_folding = folding;
  p.enable_presolve = presolve;
  p.enable_structure_specific_rules = new_rules;
  if (std::isfinite(time_limit) && time_limit > 0.0) p.max_time = time_limit;
  h.folding_summary = gp::preprocess(h.original_lp, p);
  gp::presolve::GpuPresolveSummary& s = h.folding_summary.presolve;
  h.presolve_seconds = s.elapsed_seconds;
  std::cout << h.folding_summary.folding.profile.total_seconds << " "
            << h.folding_summary.original_rows << " " << h.folding_summary.folded_cols << " "
            << h.folding_summary.folding.applied << " " << h.folding_summary.folding.owns_folded_lp << " "
            << s.reduced_rows << " " << s.reduced_cols << " " << s.record.obj_constant_new << "\n";
  if (s.has_infeasible || s.has_unbounded) return false;
  gp::presolve::export_recovery_record(h.folding_summary.presolve.record);
  const std::size_t fixed = s.record.fixed_idx.size() + s.record.tape.types.size();
  const gp::presolve::LPInfoGpu& reduced = s.reduced_lp;
  (void)reduced.A.nnz;
  (void)fixed;
  (void)h.folding_summary.workspace_setup_seconds;
  (void)h.folding_summary.unsupported_stages;
  return true;
}

gp::Solution finish(Handle& h, double* x, double* y, double* z) {
  gp::Solution sol = gp::postprocess(x, y, z, h.folding_summary, h.original_lp);
  gp::free_summary(h.folding_summary);
  return sol;
}
"""


def test_adapter_in_reference_style_compiles(tmp_path):
    src = tmp_path / "adapter.cpp"
    src.write_text(ADAPTER)
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-I", os.path.join(REPO, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_example_caller_runs():
    exe = os.path.join(PKG, "presolve_example")
    assert os.path.exists(exe), "build the package first"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    assert "x = (2, 3, 1, 1.5), y = (1, -2, 0.5), objective 1: KKT conditions hold" in r.stdout
    assert "unchanged run borrows the input LP" in r.stdout


@pytest.mark.gpu
def test_mps_tool_reads_and_presolves(tmp_path):
    import numpy as np

    import gpupre as G
    import mps_texts as T

    exe = os.path.join(PKG, "mps_presolve")
    assert os.path.exists(exe), "build the package first"
    r = subprocess.run([exe, os.path.join(REPO, "tests", "golden", "model.mps")], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr, flush=True)
    out = dict(line.split(" ", 1) for line in r.stdout.splitlines())
    assert r.returncode == 0 and out["status"] == "ok"
    assert (out["original_rows"], out["original_cols"], out["original_nnz"]) == ("2", "2", "4")

    text, _ = T.write_integer_lp(np.random.default_rng(21), 2000, 3000, 14_000)
    path = tmp_path / "intlp.mps"
    path.write_bytes(text)
    r = subprocess.run([exe, str(path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    out = dict(line.split(" ", 1) for line in r.stdout.splitlines())
    with G.read_mps(str(path)) as model, model.upload() as d:
        s = G.preprocess(d, G.default_params(enable_folding=0, enable_structure_specific_rules=0))
        try:
            want = {"original_rows": model.m, "original_cols": model.n, "reduced_rows": s.reduced_rows,
                    "reduced_cols": s.reduced_cols, "reduced_nnz": s.reduced_nnz, "iterations": s.iterations}
            status = "infeasible" if s.has_infeasible else "unbounded" if s.has_unbounded else "ok"
        finally:
            s.free()
    assert out["status"] == status and r.returncode == (0 if status == "ok" else 1)
    for k, v in want.items():
        assert out[k] == str(v), k
