"""The column colouring against the reference's own: golden colour vectors
written by tests/golden/make_colour_golden.py, which ran the reference's
colour_sparsity (modax/_sparsity.py:182-210) on each pattern. Pins the exact
vector, which the reference's tests leave open (they check counts)."""

import json
import os

import pytest

from menodax_amd._sparsity import colour_sparsity

_GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "colour_golden.json")
_CASES = json.load(open(_GOLDEN))["cases"]


@pytest.mark.parametrize("case", _CASES, ids=[c["name"] for c in _CASES])
def test_colouring_equals_the_reference(case):
    got = colour_sparsity(tuple(tuple(r) for r in case["pattern"]))
    assert list(got.colour) == case["colour"]
    assert got.n_colours == case["n_colours"]
