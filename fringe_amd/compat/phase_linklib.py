"""Top-level alias so drivers written against the reference bindings
(`import phase_linklib`) load this library: put fringe_amd/compat on PYTHONPATH."""
from fringe_amd.phase_linklib import *  # noqa: F401,F403
