"""The HIP rules against the reference's per-rule known answers."""
import pytest

from backends import check_kat, load_kats, run_kat

CASES = load_kats()


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_gpu_reference_kat(case):
    print(f"[kat] {case['name']}", flush=True)
    b, states = run_kat(case, "gpu")
    try:
        check_kat(case, b, states)
    finally:
        b.free()


from kat_programs import PROGRAMS  # noqa: E402


@pytest.mark.gpu
@pytest.mark.parametrize("name,program", PROGRAMS, ids=[p[0] for p in PROGRAMS])
def test_gpu_reference_program(name, program):
    print(f"[kat program] {name}", flush=True)
    program("gpu")
