"""The oracle's Local_Alignment restatement against the reference's own
Local_Alignment (golden vectors made by oracle/_ref/la_vectors)."""

import pytest

import la_golden
import oracle_lib


@pytest.mark.parametrize("name", la_golden.SETS)
def test_oracle_matches_reference_vectors(name):
    ts, want = la_golden.load(name)
    got, status = oracle_lib.run_tasks(ts)
    bad = la_golden.diff(got, want, status)
    assert not bad, "\n".join(bad)


def test_oracle_catches_planted_error():
    """A plausible slip (off-by-one in one trace value) must be detected."""
    ts, want = la_golden.load("exact")
    got, status = oracle_lib.run_tasks(ts)
    f, at, bt = got[3]
    at = at.copy()
    at[1] += 1
    got[3] = (f, at, bt)
    assert la_golden.diff(got, want, status)
