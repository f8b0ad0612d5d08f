"""The entry point's smoke run (one small fused run of two instances and one small
differentiable run, both against the oracle) as a GPU test, so that the test suite covers
what the entry point checks."""
from __future__ import annotations

import pytest

pytestmark = pytest.mark.gpu


def test_entry_point_smoke(require_gpu):
    import __graft_entry__
    __graft_entry__.smoke()
