"""minimc_amd — MI355X-native particle-history tracker for minimc problems.

The compute path is the HIP library minimc_amd/_build/libminimc_amd.so
(C ABI: include/minimc_amd.h).  Importing the package does not load it;
`minimc_amd.tracker.lib()` does, and raises if it was not built.
"""
from . import abi  # noqa: F401
from .world import Problem, load_xml  # noqa: F401

__all__ = ["abi", "Problem", "load_xml"]
