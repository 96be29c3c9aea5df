"""Display of tensors (format/format_tensor.rs:25-226): the host-side layout walk against the
reference's own printed outputs (doc comments in rstsr-core/tests/doc_draft) and against layout
invariants (elision thresholds MIN_PRINT 3 / MAX_PRINT 8, zero-size axes, 0-d). No GPU."""
import json
import os

import numpy as np
import pytest

from rstsr_amd import format_tensor

HERE = os.path.dirname(os.path.abspath(__file__))
REF = json.load(open(os.path.join(HERE, "golden", "reference_cases.json")))["display"]


@pytest.mark.parametrize("case", REF["cases"], ids=lambda c: c["out"][:20])
def test_reference_display_outputs(case):
    v = np.array(case["data"])
    assert format_tensor(v, case["shape"], case["stride"], case["offset"]) == case["out"]


def test_elision_thresholds():
    v = np.arange(100)
    # last axis: up to max(8, 7) = 8 printed in full, 9 elided to 3 ... 3
    assert format_tensor(v, [8], [1], 0) == "[ 0 1 2 3 4 5 6 7]"
    assert format_tensor(v, [9], [1], 0) == "[ 0 1 2 ... 6 7 8]"
    # outer axes: fewer than 8 rows printed in full, 8 rows elided
    s7 = format_tensor(v, [7, 2], [2, 1], 0)
    assert s7.count("\n") == 6 and "..." not in s7
    s8 = format_tensor(v, [8, 2], [2, 1], 0)
    assert s8 == "[[ 0 1]\n [ 2 3]\n [ 4 5]\n ...\n [ 10 11]\n [ 12 13]\n [ 14 15]]"


def test_degenerate_layouts():
    v = np.arange(16)
    assert format_tensor(v, [2, 0, 4], [0, 4, 1], 0) == "[]"
    assert format_tensor(v, [2, 4, 0], [0, 1, 1], 0) == "[]"
    assert format_tensor(v, [], [], 5) == "5"
    assert format_tensor(np.array([1.0, 0.5, -0.0, np.nan, 1e21, -np.inf]), [6], [1], 0) == \
        "[ 1 0.5 -0 NaN 1000000000000000000000 -inf]"
    assert format_tensor(np.array([1 + 2j, 1 - 2j]), [2], [1], 0) == "[ 1+2i 1-2i]"
