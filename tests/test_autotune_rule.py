"""CPU test of the autotuner's result-equality rule: the first candidate's
digest is the reference, a later candidate with another digest is an
error, and digests depend only on the result values."""
from __future__ import annotations

import numpy as np
import pytest

import autotune


def test_first_digest_is_the_reference_and_a_mismatch_raises():
    ref = autotune.check_same(None, "aa", "first")
    assert ref == "aa"
    assert autotune.check_same(ref, "aa", "second") == "aa"
    with pytest.raises(autotune.CandidateMismatch, match="third"):
        autotune.check_same(ref, "bb", "third")


def test_digest_depends_on_values_only():
    a = np.array([1, 2, 3], dtype=np.uint64)
    b = np.array([4, -5], dtype=np.int64)
    assert autotune.digest(a, b) == autotune.digest(a.copy(), b.copy())
    assert autotune.digest(a, b) != autotune.digest(a, b + 1)
    assert autotune.digest(a[::-1].copy(), b) != autotune.digest(a, b)
