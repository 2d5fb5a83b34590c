"""The CPU oracle against the golden vectors of the real OpenPhase code (bit for bit)."""
import numpy as np
import pytest

from _golden import CASES, Case
import oracle as O


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference_bitwise(name):
    case = Case(name)
    orc = case.setup(O.Oracle(case.cfg))
    for s, got, vol, vr in case.run(orc):
        exp, evol, evr = case.expect[s]
        assert np.array_equal(got.count, exp.count), f"{name} step {s}: entry counts differ"
        assert np.array_equal(got.index, exp.index), f"{name} step {s}: active index lists differ"
        assert np.array_equal(got.value, exp.value), (
            f"{name} step {s}: values differ, max {np.abs(got.value - exp.value).max()}")
        assert np.array_equal(vol, evol), f"{name} step {s}: grain volumes differ"
        assert np.array_equal(vr, evr), f"{name} step {s}: volume ratios differ"
