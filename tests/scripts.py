"""Workload scripts shared by the tests (defined with the package)."""
from map_amd.workloads import *  # noqa: F401,F403
from map_amd.workloads import PI, LIFE_MASK  # noqa: F401
