"""The oracle restatement (oracle/_bin/daligner_oracle) writes the same .las
bytes as the compiled reference daligner for every golden case."""

import json
import os
import tempfile

import pytest

import las_cases
import oracle_lib

GOLD = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "las", "goldens.json")))


@pytest.mark.parametrize("case", sorted(las_cases.CASES))
def test_oracle_matches_reference_las(case):
    if case in las_cases.LARGE and not os.environ.get("MI_SLOW"):
        pytest.skip("large workload: set MI_SLOW=1 (minutes on one core)")
    oracle_lib.ensure_built()
    with tempfile.TemporaryDirectory() as d:
        las_cases.make_db(case, d)
        got = las_cases.run(oracle_lib.ORACLE_CLI, case, d)
    assert got == GOLD[case]["las"]
