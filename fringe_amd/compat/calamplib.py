"""Top-level alias so drivers written against the reference bindings
(`import calamplib`) load this library: put fringe_amd/compat on PYTHONPATH."""
from fringe_amd.calamplib import *  # noqa: F401,F403
