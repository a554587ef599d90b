"""The CPU oracle of the DFSPH step — TEST INFRASTRUCTURE ONLY (see oracle.py)."""
from .oracle import FIELDS, OracleSim, build, lib, max_threads, set_threads

__all__ = ["FIELDS", "OracleSim", "build", "lib", "max_threads", "set_threads"]
