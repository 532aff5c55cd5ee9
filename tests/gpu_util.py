"""Helpers shared by the GPU tests."""
import pytest


def require_gpu():
    from minimc_amd import tracker
    try:
        n = tracker.device_count()
    except RuntimeError as e:  # library missing: a GPU test must fail loudly
        raise AssertionError(str(e))
    if n < 1:
        pytest.fail("no HIP device visible")
