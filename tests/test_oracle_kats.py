"""The CPU restatement against the reference's own per-rule known answers."""
import pytest

from backends import check_kat, load_kats, run_kat

CASES = load_kats()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_reference_kat(case):
    b, states = run_kat(case, "oracle")
    check_kat(case, b, states)


from kat_programs import PROGRAMS  # noqa: E402


@pytest.mark.parametrize("name,program", PROGRAMS, ids=[p[0] for p in PROGRAMS])
def test_oracle_reference_program(name, program):
    program("oracle")
