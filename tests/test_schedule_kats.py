"""The reference's scheduler test cases (test_deferred_schedule.cpp) on the
CPU restatement and on the HIP engine."""
import pytest

from schedule_programs import PROGRAMS


@pytest.mark.parametrize("name,program", PROGRAMS, ids=[p[0] for p in PROGRAMS])
def test_oracle_schedule(name, program):
    program("oracle")


@pytest.mark.gpu
@pytest.mark.parametrize("name,program", PROGRAMS, ids=[p[0] for p in PROGRAMS])
def test_gpu_schedule(name, program):
    print(f"[schedule] {name}", flush=True)
    program("gpu")
