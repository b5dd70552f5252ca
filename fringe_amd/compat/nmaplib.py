"""Top-level alias so drivers written against the reference bindings
(`import nmaplib`) load this library: put fringe_amd/compat on PYTHONPATH."""
from fringe_amd.nmaplib import *  # noqa: F401,F403
