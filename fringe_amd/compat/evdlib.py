"""Top-level alias so drivers written against the reference bindings
(`import evdlib`) load this library: put fringe_amd/compat on PYTHONPATH."""
from fringe_amd.evdlib import *  # noqa: F401,F403
