"""The C++ front end with the reference's API names (include/gpu_preprocessor/
gpu_preprocessor.hpp): it compiles as plain host C++ (layout checks against
the C ABI are static_asserts), and on a GPU the example caller
(examples/presolve_example.cpp) runs preprocess / postprocess / free_summary
on an LP that presolve solves completely and checks the recovered point's
KKT conditions."""
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
        "  return p.max_iters == 512 && lp.A.rows == 2 ? 0 : 1;\n"
        "}\n")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(REPO, "include"), str(src)],
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
